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


def _plugin_deps(d, seen=None):
    """A plug-in directory plus the directories its headers include from
    (e.g. a sibling plug-in's helper header)."""
    import re
    seen = set() if seen is None else seen
    d = os.path.abspath(d)
    if d in seen or not os.path.isdir(d):
        return seen
    seen.add(d)
    for f in os.listdir(d):
        fp = os.path.join(d, f)
        if not os.path.isfile(fp):
            continue
        for inc in re.findall(r'#include\s+"([^"]+)"', open(fp, errors="replace").read()):
            tgt = os.path.dirname(os.path.abspath(os.path.join(d, inc)))
            if tgt != d:
                _plugin_deps(tgt, seen)
    return seen


def _compile(src, obj, deps, extra=()):
    stamp = obj + ".sha"
    key = _deps_hash([src] + list(deps)) + ARCH + " ".join(extra)
    if os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key:
        return obj
    _run([HIPCC, "--offload-arch=" + ARCH] + COMMON + list(extra) + ["-c", src, "-o", obj])
    with open(stamp, "w") as f:
        f.write(key)
    return obj


# Batch units: sets of plug-ins compiled into one persistent kernel per
# method (cdm_batch_levmar / cdm_batch_mcmc).  The default set is the batched
# workload of tensor, two-fan Bingham-NODDI and three-stick voxels.
DEFAULT_BATCHES = {"tensor_bingham2_sticks3": ["DiffusionTensor", "NODDI_Bingham_2Fans", "Ball_3_Sticks"]}

PLUGIN_MACROS = ["MODEL_H_INCLUDED", "NPARAMS", "NCFP", "NFIXP", "CDM_LM_JACOBIAN_DOUBLE", "CDM_VOXEL_CACHE",
                 "CDM_VOXEL_CACHE_DEPENDS", "CDM_MEAS_CACHE", "CDM_NUMERICAL_MASK", "CDM_MEAS_VOXEL_CACHE",
                 "CDM_MEAS_VOXEL_CACHE_DEPENDS", "CDM_VOXEL_CACHE_JOBS", "CDM_VOXEL_CACHE_JOBS_DEPENDS",
                 "CDM_MEAS_VOXEL_CACHE_FILL", "CDM_MEAS_VOXEL_CACHE_PARTIAL", "CDM_VOXEL_CACHE_PARTIAL", "CDM_SIGNAL_PARTIALS", "CDM_WATSON_REFERENCE_FORMS",
                 "CDM_LM_JACOBIAN_COLUMNS", "CDM_LM_ROW_PASSES"]


def batch_unit_source(name, model_dirs):
    """Source of one batch unit: each plug-in in its own namespace with the
    engine bodies, then the persistent kernels and the host-side tables."""
    L = [1, 4, 16, 64]
    out = ["// generated: batch unit %s (%s)" % (name, ", ".join(m for m, _ in model_dirs)),
           '#include "abi/functions_gpu.h"', '#include "engine/philox.h"', '#include "engine/wave.h"',
           '#include "engine/packed.h"', '#include "host/model_ops.h"', "#include <algorithm>", "#include <climits>",
           "#include <cstring>", ""]
    for k, (m, d) in enumerate(model_dirs):
        out += ["#undef %s" % x for x in PLUGIN_MACROS]
        out += ["namespace cdm_batch_%s_m%d {" % (name, k),
                '#include "%s/modelparameters.h"' % d, '#include "%s/modelparameters.cc"' % d,
                '#include "abi/macro_numerical.h"', '#include "%s/modelfunctions.h"' % d,
                '#include "engine/engine.h"', '#include "engine/args.inl"',
                "template <int L>",
                "__device__ __noinline__ void batch_lm(const void* a, int marquardt, int blk) {",
                "  cdm_levmar_body<L>(*static_cast<const cdm_lm_args*>(a), marquardt, blk);", "}",
                "template <int L, bool RICIAN>",
                "__device__ __noinline__ void batch_mcmc(const void* a, int blk) {",
                "  cdm_mcmc_body<L, RICIAN>(*static_cast<const cdm_mcmc_args*>(a), blk);", "}",
                "}  // namespace", ""]
    ns = lambda k: "cdm_batch_%s_m%d" % (name, k)
    sizes = ", ".join("sizeof(%s::cdm_lm_args), sizeof(%s::cdm_mcmc_args)" % (ns(k), ns(k))
                      for k in range(len(model_dirs)))
    out += ["namespace cdm_batch_%s {" % name,
            "constexpr size_t kSizes[] = {%s};" % sizes,
            "constexpr size_t arg_bytes() { size_t m = 0; for (size_t s : kSizes) m = s > m ? s : m; return m; }",
            "struct DevItem {", "  int kase, first_tile, ntiles, pad;", "  alignas(16) unsigned char args[arg_bytes()];",
            "};", "",
            "template <int MODE>",
            "__global__ void __launch_bounds__(256) batch_kernel(const DevItem* __restrict__ items, int nitems,",
            "                                                     int total_tiles, int* queue, int marquardt) {",
            "  __shared__ int s_tile;", "  for (;;) {", "    __syncthreads();",
            "    if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1);", "    __syncthreads();",
            "    const int tile = s_tile;", "    if (tile >= total_tiles) break;", "    int it = 0;",
            "    while (it + 1 < nitems && tile >= items[it + 1].first_tile) ++it;",
            "    const DevItem* d = &items[it];", "    const int blk = tile - d->first_tile;", "    switch (d->kase) {"]
    for k in range(len(model_dirs)):
        for li, l in enumerate(L):
            out.append("      case %d: if constexpr (MODE == cdm::kBatchLevMar) %s::batch_lm<%d>(d->args, marquardt, blk); "
                       "else %s::batch_mcmc<%d, MODE == cdm::kBatchMcmcRician>(d->args, blk); break;" %
                       (4 * k + li, ns(k), l, ns(k), l))
    out += ["      default: break;", "    }", "  }", "}", "",
            "// one persistent kernel per mode: LM, MCMC, MCMC with the Rician likelihood",
            "static auto kernel_of(int mode) {",
            "  return mode == cdm::kBatchLevMar ? batch_kernel<cdm::kBatchLevMar>",
            "         : mode == cdm::kBatchMcmc  ? batch_kernel<cdm::kBatchMcmc>",
            "                                    : batch_kernel<cdm::kBatchMcmcRician>;", "}", ""]
    out += ["// the persistent MCMC kernel keeps cache rows in LDS up to 80 KB a block, as a",
            "// two-blocks-per-CU budget would",
            "constexpr int kBatchMcmcBudget = 2;",
            "static int prepare(int mode, int nitems, const cdm::BatchLaunchItem* in, void* host_out, int* total_tiles,"]
    out += ["                   size_t* lds_bytes) {",
            "  DevItem* out = static_cast<DevItem*>(host_out);", "  int tiles = 0;", "  size_t lds = 0;",
            "  for (int i = 0; i < nitems; ++i) {", "    const cdm::BatchLaunchItem& b = in[i];",
            "    if (b.model < 0 || b.model >= %d || b.g.nvox <= 0) return -1;" % len(model_dirs),
            "    DevItem& d = out[i];", "    std::memset(&d, 0, sizeof(d));",
            "    int L = b.L;", "    size_t need = 0;", "    switch (b.model) {"]
    for k in range(len(model_dirs)):
        out += ["      case %d: {" % k,
                "        if (b.g.cfp_ts > %s::cdm_cfp_max || b.g.fixp_ts > %s::cdm_fixp_max) return -1;" % (ns(k), ns(k)),
                "        // LM groups of 64 lanes past the LDS fall back to 16, as the per-model launch",
                "        if (mode == cdm::kBatchLevMar) L = %s::op_lm_lanes(b.g, L);" % ns(k),
                "        need = mode != cdm::kBatchLevMar ? %s::op_lds_mcmc(b.g, L, kBatchMcmcBudget) : %s::op_lds_lm(b.g, L);" % (ns(k), ns(k)),
                "        if (need > %s::cdm_lds_max) return -1;" % ns(k),
                "        if (mode == cdm::kBatchLevMar) {",
                "          const auto a = %s::op_lm_args(b.g, b.params, b.max_iters, b.c, b.iters, b.cost);" % ns(k),
                "          std::memcpy(d.args, &a, sizeof(a));", "        } else {",
                "          const auto a = %s::op_mcmc_args(b.g, L, kBatchMcmcBudget, b.mc, b.c);" % ns(k),
                "          std::memcpy(d.args, &a, sizeof(a));", "        }", "      } break;"]
    out += ["      default: return -1;", "    }",
            "    const int li = L == 1 ? 0 : L == 4 ? 1 : L == 16 ? 2 : L == 64 ? 3 : -1;",
            "    if (li < 0) return -1;",
            "    d.kase = 4 * b.model + li;", "    d.first_tile = tiles;", "    const int per = 256 / L;",
            "    d.ntiles = (b.g.nvox + per - 1) / per;", "    tiles += d.ntiles;",
            "    lds = std::max(lds, need);", "  }",
            "  *total_tiles = tiles;", "  *lds_bytes = lds;", "  return 0;", "}", "",
            "static hipError_t grid(int mode, int total_tiles, size_t lds, int* out) {",
            "  auto kern = kernel_of(mode);",
            "  int dev = 0, cus = 0, per_cu = 0;", "  hipError_t e = hipGetDevice(&dev);",
            "  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);",
            "  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds);",
            "  if (e != hipSuccess) return e;",
            "  *out = std::max(1, std::min(total_tiles, std::max(1, cus) * std::max(1, per_cu)));",
            "  return hipSuccess;", "}", "",
            "static hipError_t launch(int mode, const void* dev_items, int nitems, int total_tiles, int* queue,",
            "                         size_t lds, int marquardt, int grid, hipStream_t s) {",
            "  if (total_tiles <= 0) return hipSuccess;",
            "  auto kern = kernel_of(mode);",
            "  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, static_cast<const DevItem*>(dev_items), nitems,",
            "                     total_tiles, queue, marquardt);",
            "  return hipGetLastError();", "}", "",
            "#ifndef __HIP_DEVICE_COMPILE__  // host tables (constant initialisers would be copied to the device)",
            "static const char* const kNames[] = {%s};" % ", ".join('"%s"' % m for m, _ in model_dirs),
            "extern const cdm::BatchOps ops = {\"%s\", %d, kNames, sizeof(DevItem), &prepare, &grid, &launch};" %
            (name, len(model_dirs)),
            "#endif", "}  // namespace", ""]
    return "\n".join(out)


def build(models=None, extra_dirs=(), force=False, jobs=None, batches=None, out_dir=None):
    """Compile the engine library.  models: shipped plug-ins (None: all, []:
    none); extra_dirs: further plug-in directories in the reference's
    format, each compiled unmodified under its directory's name.  out_dir:
    build tree and library under out_dir (out_dir/_build,
    out_dir/lib/libcudimot.so) instead of the package's own, with no host
    driver or FLOP-model library; a scratch build never replaces the
    package's library."""
    build_dir, lib_dir = (BUILD, LIBDIR) if out_dir is None else (os.path.join(out_dir, "_build"),
                                                                    os.path.join(out_dir, "lib"))
    lib_path = os.path.join(lib_dir, "libcudimot.so")
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(lib_dir, exist_ok=True)
    dirs = {}
    for m in (shipped_models() if models is None else models):
        dirs[m] = os.path.join(MODELS, m)
    for d in extra_dirs:
        d = os.path.abspath(d)
        dirs[os.path.basename(d.rstrip("/"))] = d
    template = open(os.path.join(CSRC, "engine", "model_unit.hip.in")).read()
    engine_deps = [os.path.join(CSRC, "engine"), os.path.join(CSRC, "abi"), os.path.join(CSRC, "host", "model_ops.h")]
    jobs_list = []
    for name, d in sorted(dirs.items()):
        src = os.path.join(build_dir, "unit_%s.hip" % name)
        text = template.replace("@NAME@", name).replace("@DIR@", d)
        if not os.path.exists(src) or open(src).read() != text:
            with open(src, "w") as f:
                f.write(text)
        jobs_list.append((src, os.path.join(build_dir, "unit_%s.o" % name), engine_deps + sorted(_plugin_deps(d))))
    batches = DEFAULT_BATCHES if batches is None else batches
    batches = {b: ms for b, ms in batches.items() if all(m in dirs for m in ms)}
    for bname, ms in sorted(batches.items()):
        src = os.path.join(build_dir, "batch_%s.hip" % bname)
        text = batch_unit_source(bname, [(m, dirs[m]) for m in ms])
        if not os.path.exists(src) or open(src).read() != text:
            with open(src, "w") as f:
                f.write(text)
        bdeps = set()
        for m in ms:
            _plugin_deps(dirs[m], bdeps)
        jobs_list.append((src, os.path.join(build_dir, "batch_%s.o" % bname), engine_deps + sorted(bdeps)))
    reg = os.path.join(build_dir, "registry.cpp")
    lines = ['#include "host/model_ops.h"', "#include <cstring>"]
    for name in sorted(dirs):
        lines.append("namespace cdm_model_%s { extern const cdm::ModelOps ops; }" % name)
    for bname in sorted(batches):
        lines.append("namespace cdm_batch_%s { extern const cdm::BatchOps ops; }" % bname)
    lines.append("namespace cdm {")
    lines.append("static const BatchOps* const kBatches[] = {%s};" %
                 ", ".join(["&cdm_batch_%s::ops" % b for b in sorted(batches)] + ["nullptr"]))
    lines.append("int batch_count() { return %d; }" % len(batches))
    lines.append("const BatchOps* batch_at(int i) { return kBatches[i]; }")
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
    jobs_list.append((reg, os.path.join(build_dir, "registry.o"), host_deps))
    for f in ("capi.cpp", "spec.cpp", "noalloc.hip", "fill.hip"):
        jobs_list.append((os.path.join(CSRC, "host", f), os.path.join(build_dir, f.split(".")[0] + ".o"), host_deps))
    jobs_list.append((os.path.join(CSRC, "mlp", "mlp.hip"), os.path.join(build_dir, "mlp.o"),
                      host_deps + [os.path.join(CSRC, "mlp"), os.path.join(CSRC, "engine", "philox.h")]))
    jobs_list.append((os.path.join(CSRC, "init", "orientation.hip"), os.path.join(build_dir, "orientation.o"),
                      host_deps + [os.path.join(CSRC, "init")]))
    jobs_list.append((os.path.join(CSRC, "finish", "derived.hip"), os.path.join(build_dir, "derived.o"),
                      host_deps + [os.path.join(CSRC, "finish")]))
    if force:
        for _, o, _ in jobs_list:
            if os.path.exists(o + ".sha"):
                os.remove(o + ".sha")
    n = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs_list))
    tmp = lib_path + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, lib_path)
    if out_dir is None:
        build_cli()
        build_opcount(models)
    return lib_path


BINDIR = os.path.join(PKG, "bin")
CLI = os.path.join(BINDIR, "cudimot")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_cli():
    """Host driver ``bin/cudimot`` (csrc/cli): the reference's split_parts,
    fit and merge_parts programs over the C-ABI.  Links the engine library
    and zlib; symlinks named split_parts_<Model>, <Model> and
    merge_parts_<Model> select the step as the reference's binaries do."""
    os.makedirs(BINDIR, exist_ok=True)
    inc = os.path.join(PKG, "..", "include")
    deps = [os.path.join(CSRC, "cli"), os.path.join(inc, "cudimot_amd.h")]
    flags = ["-O2", "-std=c++17", "-I", CSRC, "-I", inc, "-I", os.path.join(ROCM, "include"), "-D__HIP_PLATFORM_AMD__"]
    objs = []
    for f in ("nifti_io.cpp", "parts.cpp", "options.cpp", "driver.cpp"):
        src = os.path.join(CSRC, "cli", f)
        obj = os.path.join(BUILD, "cli_" + f[:-4] + ".o")
        stamp = obj + ".sha"
        key = _deps_hash([src] + deps)
        if not (os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key):
            _run(["g++"] + flags + ["-c", src, "-o", obj])
            with open(stamp, "w") as fh:
                fh.write(key)
        objs.append(obj)
    tmp = CLI + ".tmp"
    _run(["g++", "-o", tmp] + objs + ["-L", LIBDIR, "-lcudimot", "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-lz",
                                      "-Wl,-rpath,$ORIGIN/../lib:" + os.path.join(ROCM, "lib")])
    os.replace(tmp, CLI)
    return CLI


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
        jobs_list.append((src, os.path.join(BUILD, "opc_%s.o" % name), sorted(_plugin_deps(d)) + [os.path.join(CSRC, "tools"),
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
    ap.add_argument("--models", default="", help="comma-separated shipped plug-ins (default: all; 'none': none)")
    ap.add_argument("--extra-model-dir", action="append", default=[],
                    help="a plug-in directory in the reference's format to add, compiled unmodified")
    ap.add_argument("--out", default=None,
                    help="build into this directory (OUT/lib/libcudimot.so) instead of the package's library")
    ap.add_argument("--batch", action="append", default=[],
                    help="a batch unit NAME=ModelA,ModelB,... (default: %s)" % DEFAULT_BATCHES)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args(argv)
    models = [] if a.models == "none" else ([m for m in a.models.split(",") if m] or None)
    batches = dict((b.split("=", 1)[0], b.split("=", 1)[1].split(",")) for b in a.batch) if a.batch else None
    print(build(models, a.extra_model_dir, a.force, batches=batches, out_dir=a.out))
    if a.out is None:
        print(OPC_LIB)


if __name__ == "__main__":
    sys.exit(main())
