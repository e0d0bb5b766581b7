"""ORACLE build recipe (test infrastructure only; never part of the product).

Builds two CPU libraries with g++:

* ``oracle/_build/liboracle.so`` -- the FP64 restatement of the reference's
  per-voxel GridSearch / Levenberg-Marquardt / MCMC / predicted-signal /
  BIC-AIC loops (oracle_engine.h), instantiated for every plug-in under
  ``cudimot_amd/mymodels`` with the host plug-in ABI (host_abi.h).
* ``oracle/_ref/libref_models.so`` -- only when the reference checkout is
  present: the same engine instantiated with the reference's own plug-in
  headers, compiled from where they lie under /root/reference/mymodels.  Its
  outputs pin this project's plug-ins and feed the golden fixtures.
* ``oracle/_ref/dropin/lib/libcudimot.so`` -- only when the reference
  checkout is present: the GPU engine library built by cudimot_amd.build
  over reference plug-in directories, unmodified (the plug-in drop-in the
  tests run against the oracle over the same headers).

Run: ``python oracle/build_oracle.py``.  Outputs go only to oracle/_build and
oracle/_ref (both git-ignored).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUR_MODELS = os.path.join(REPO, "cudimot_amd", "mymodels")
REF_MODELS = "/root/reference/mymodels"
REF_SET = ["Ball_1_Stick", "Ball_2_Sticks", "DiffusionTensor", "NODDI_Watson", "NODDI_Bingham",
           "NODDI_Bingham_2Fans", "toymodel"]
CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c++17", "-w"]

UNIT = """// generated
#include "host_abi.h"
#include "philox_ref.h"
#include <omp.h>
#include <stdint.h>
#undef MODEL_H_INCLUDED
#undef NPARAMS
#undef NCFP
#undef NFIXP
namespace {ns} {{
#include "{d}/modelparameters.h"
#include "{d}/modelparameters.cc"
#include "numerical_host.h"
#include "{d}/modelfunctions.h"
#include "oracle_engine.h"
#define ORC_PREFIX {prefix}
#include "oracle_exports.inc"
#undef ORC_PREFIX
}}
"""


def _compile(src, obj):
    cmd = [CXX] + FLAGS + ["-I", HERE, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle compile failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    return obj


def _build_lib(outdir, libname, units):
    os.makedirs(outdir, exist_ok=True)
    srcs = []
    for name, (ns, d, prefix) in units.items():
        src = os.path.join(outdir, "unit_%s.cpp" % name)
        text = UNIT.format(ns=ns, d=d, prefix=prefix)
        if not os.path.exists(src) or open(src).read() != text:
            with open(src, "w") as f:
                f.write(text)
        srcs.append(src)
    srcs.append(os.path.join(HERE, "orc_common.cpp"))
    objs = [os.path.join(outdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    lib = os.path.join(outdir, libname)
    tmp = lib + ".tmp"
    subprocess.run([CXX, "-shared", "-fopenmp", "-o", tmp] + objs, check=True)
    os.replace(tmp, lib)
    return lib


def our_models():
    return sorted(d for d in os.listdir(OUR_MODELS)
                  if os.path.exists(os.path.join(OUR_MODELS, d, "modelfunctions.h")))


def sources_digest():
    """SHA-1 over every source liboracle.so is compiled from (the oracle's
    own headers and this project's plug-ins); bind.py refuses a library
    whose stamp differs."""
    import hashlib
    h = hashlib.sha1()
    files = [os.path.join(HERE, f) for f in sorted(os.listdir(HERE))
             if f.endswith((".h", ".inc")) or f == "orc_common.cpp"]
    for root, dirs, fs in sorted(os.walk(OUR_MODELS)):
        dirs.sort()
        files += [os.path.join(root, f) for f in sorted(fs)]
    for f in files:
        h.update(os.path.relpath(f, REPO).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def build_oracle():
    units = {m: ("orc_" + m, os.path.join(OUR_MODELS, m), "orc_%s_" % m) for m in our_models()}
    digest = sources_digest()
    lib = _build_lib(os.path.join(HERE, "_build"), "liboracle.so", units)
    with open(lib + ".sha", "w") as f:
        f.write(digest + "\n")
    return lib


def build_ref():
    if not os.path.isdir(REF_MODELS):
        return None
    units = {m: ("ref_" + m, os.path.join(REF_MODELS, m), "ref_%s_" % m) for m in REF_SET
             if os.path.isdir(os.path.join(REF_MODELS, m))}
    return _build_lib(os.path.join(HERE, "_ref"), "libref_models.so", units)


# Reference plug-ins compiled unmodified through the engine's own build (the
# plug-in drop-in): a scratch engine library holding only these models.
DROPIN_SET = ["Ball_1_Stick", "NODDI_Watson"]
DROPIN_DIR = os.path.join(HERE, "_ref", "dropin")
DROPIN_LIB = os.path.join(DROPIN_DIR, "lib", "libcudimot.so")


def build_dropin():
    """The engine library (cudimot_amd.build, gfx950) over the reference's own
    plug-in directories, unmodified, into oracle/_ref/dropin; only when the
    reference checkout is present."""
    dirs = [os.path.join(REF_MODELS, m) for m in DROPIN_SET if os.path.isdir(os.path.join(REF_MODELS, m))]
    if not dirs:
        return None
    sys.path.insert(0, REPO)
    try:
        from cudimot_amd import build as engine_build
    finally:
        sys.path.pop(0)
    return engine_build.build(models=[], extra_dirs=dirs, batches={}, out_dir=DROPIN_DIR)


if __name__ == "__main__":
    print(build_oracle())
    r = build_ref()
    print(r if r else "reference checkout absent: oracle/_ref not built")
    d = build_dropin()
    print(d if d else "reference checkout absent: oracle/_ref/dropin not built")
    sys.exit(0)
