"""Build the engine library ``cudimot_amd/lib/libcudimot.so`` with hipcc.

One translation unit per model plug-in (``mymodels/<name>/`` in the
reference's plug-in format) is generated from
``csrc/engine/model_unit.hip.in`` and compiled for gfx950; the host C-ABI
(``csrc/host``) and a generated registry are linked in.  The reference
builds one binary per model (``modelname=X make``); here every selected
plug-in is compiled into one library and chosen by name at run time.

    python -m cudimot_amd.build [--models A,B] [--extra-model-dir DIR ...]
"""
import argparse
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libcudimot.so")
MODELS = os.path.join(PKG, "mymodels")
ARCH = os.environ.get("CDM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", CSRC, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def shipped_models():
    return sorted(d for d in os.listdir(MODELS) if os.path.exists(os.path.join(MODELS, d, "modelfunctions.h")))


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), err[-6000:]))


def _deps_hash(paths):
    h = hashlib.sha1()
    for p in paths:
        if os.path.isdir(p):
            for root, _, files in sorted(os.walk(p)):
                for f in sorted(files):
                    fp = os.path.join(root, f)
                    h.update(fp.encode())
                    h.update(open(fp, "rb").read())
        else:
            h.update(p.encode())
            h.update(open(p, "rb").read())
    return h.hexdigest()


def _compile(src, obj, deps, extra=()):
    stamp = obj + ".sha"
    key = _deps_hash([src] + list(deps)) + ARCH + " ".join(extra)
    if os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key:
        return obj
    _run([HIPCC, "--offload-arch=" + ARCH] + COMMON + list(extra) + ["-c", src, "-o", obj])
    with open(stamp, "w") as f:
        f.write(key)
    return obj


def build(models=None, extra_dirs=(), force=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    dirs = {}
    for m in (models or shipped_models()):
        dirs[m] = os.path.join(MODELS, m)
    for d in extra_dirs:
        d = os.path.abspath(d)
        dirs[os.path.basename(d.rstrip("/"))] = d
    template = open(os.path.join(CSRC, "engine", "model_unit.hip.in")).read()
    engine_deps = [os.path.join(CSRC, "engine"), os.path.join(CSRC, "abi"), os.path.join(CSRC, "host", "model_ops.h")]
    jobs_list = []
    for name, d in sorted(dirs.items()):
        src = os.path.join(BUILD, "unit_%s.hip" % name)
        text = template.replace("@NAME@", name).replace("@DIR@", d)
        if not os.path.exists(src) or open(src).read() != text:
            with open(src, "w") as f:
                f.write(text)
        jobs_list.append((src, os.path.join(BUILD, "unit_%s.o" % name), engine_deps + [d]))
    reg = os.path.join(BUILD, "registry.cpp")
    lines = ['#include "host/model_ops.h"', "#include <cstring>"]
    for name in sorted(dirs):
        lines.append("namespace cdm_model_%s { extern const cdm::ModelOps ops; }" % name)
    lines.append("namespace cdm {")
    lines.append("static const ModelOps* const kModels[] = {%s};" %
                 ", ".join("&cdm_model_%s::ops" % n for n in sorted(dirs)))
    lines.append("int model_count() { return (int)(sizeof(kModels) / sizeof(kModels[0])); }")
    lines.append("const ModelOps* model_at(int i) { return kModels[i]; }")
    lines.append("const ModelOps* find_model(const char* n) { for (const ModelOps* m : kModels) "
                 "if (std::strcmp(m->name, n) == 0) return m; return nullptr; }")
    lines.append("}")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(reg) or open(reg).read() != text:
        with open(reg, "w") as f:
            f.write(text)
    host_deps = [os.path.join(CSRC, "host"), os.path.join(PKG, "..", "include", "cudimot_amd.h")]
    jobs_list.append((reg, os.path.join(BUILD, "registry.o"), host_deps))
    for f in ("capi.cpp", "spec.cpp"):
        jobs_list.append((os.path.join(CSRC, "host", f), os.path.join(BUILD, f[:-4] + ".o"), host_deps))
    jobs_list.append((os.path.join(CSRC, "mlp", "mlp.hip"), os.path.join(BUILD, "mlp.o"),
                      host_deps + [os.path.join(CSRC, "mlp"), os.path.join(CSRC, "engine", "philox.h")]))
    jobs_list.append((os.path.join(CSRC, "init", "orientation.hip"), os.path.join(BUILD, "orientation.o"),
                      host_deps + [os.path.join(CSRC, "init")]))
    if force:
        for _, o, _ in jobs_list:
            if os.path.exists(o + ".sha"):
                os.remove(o + ".sha")
    n = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs_list))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, LIB)
    return LIB


OPC_LIB = os.path.join(LIBDIR, "libcdm_opcount.so")


def build_opcount(models=None):
    """Host library of the FLOP model: every plug-in compiled with an
    operation-counting scalar (csrc/tools/opcount.h)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    template = open(os.path.join(CSRC, "tools", "opcount_unit.cpp.in")).read()
    objs = []
    jobs_list = []
    for name in (models or shipped_models()):
        d = os.path.join(MODELS, name)
        src = os.path.join(BUILD, "opc_%s.cpp" % name)
        text = template.replace("@NAME@", name).replace("@DIR@", d)
        if not os.path.exists(src) or open(src).read() != text:
            with open(src, "w") as f:
                f.write(text)
        jobs_list.append((src, os.path.join(BUILD, "opc_%s.o" % name), [d, os.path.join(CSRC, "tools"),
                                                                          os.path.join(CSRC, "abi")]))

    def comp(src, obj, deps):
        stamp = obj + ".sha"
        key = _deps_hash([src] + list(deps))
        if not (os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key):
            _run(["g++", "-O1", "-std=c++17", "-fPIC", "-w", "-I", CSRC, "-c", src, "-o", obj])
            with open(stamp, "w") as f:
                f.write(key)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda j: comp(*j), jobs_list))
    tmp = OPC_LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs)
    os.replace(tmp, OPC_LIB)
    return OPC_LIB


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--models", default="", help="comma-separated shipped plug-ins (default: all)")
    ap.add_argument("--extra-model-dir", action="append", default=[], help="a plug-in directory to add")
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args(argv)
    models = [m for m in a.models.split(",") if m] or None
    print(build(models, a.extra_model_dir, a.force))
    print(build_opcount(models))


if __name__ == "__main__":
    sys.exit(main())
