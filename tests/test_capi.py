"""The C-ABI library loads without a GPU, exports every entry point that
include/cudimot_amd.h declares, and its host-only calls work on the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from cudimot_amd import capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cudimot_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cdm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_models_compiled_in():
    assert set(capi.list_models()) >= {"Ball_1_Stick", "Ball_2_Sticks", "Ball_3_Sticks", "DiffusionTensor",
                                        "NODDI_Watson", "NODDI_Bingham", "NODDI_Bingham_2Fans", "toymodel"}


def test_unknown_model_is_an_error():
    with pytest.raises(capi.CdmError, match="not compiled"):
        capi.Model("NoSuchModel")


def test_model_info_sizes():
    m = capi.Model("NODDI_Bingham")
    assert (m.nparams, m.cfp_tsize, m.fixp_tsize) == (7, 4, 1)
    m = capi.Model("toymodel")
    assert (m.nparams, m.cfp_tsize, m.fixp_tsize) == (2, 1, 0)


def test_group_size_validation():
    m = capi.Model("Ball_1_Stick")
    m.set_group(capi.LEVMAR, 16)
    assert m.get_group(capi.LEVMAR, 64) == 16
    m.set_group(capi.LEVMAR, 0)
    assert m.get_group(capi.LEVMAR, 64) in (1, 4, 16, 64)
    with pytest.raises(capi.CdmError):
        m.set_group(capi.LEVMAR, 3)


def test_argument_errors_without_gpu():
    """Errors found on the host come back as status + message, no exit."""
    m = capi.Model("Ball_1_Stick")
    L = capi.lib()
    assert L.cdm_levmar(m.h, 10, 64, None, None, None, None, 100, 1, None, None, None) != 0
    assert "null" in L.cdm_last_error().decode()
    assert L.cdm_mcmc(m.h, 10, 64, 8, 8, 8, None, 10, 10, 25, 1, 0, 1, 0, 8, None, None) != 0
    assert "at least one sample" in L.cdm_last_error().decode()
    assert L.cdm_gridsearch(m.h, 10, 64, 8, 8, 8, None, None, None) != 0
    assert "grid" in L.cdm_last_error().decode()
    n = capi.Model("NODDI_Watson")
    assert L.cdm_levmar(n.h, 10, 99, 8, 8, 8, None, 100, 1, None, None, None) != 0
    assert "FixP" in L.cdm_last_error().decode()


def test_empty_volume_is_a_no_op_without_gpu():
    """Zero voxels: every method returns success before touching the device
    (a part of a split volume may be empty); the data pointers may be null."""
    m = capi.Model("NODDI_Bingham")
    L = capi.lib()
    cfp = 8  # never dereferenced
    assert L.cdm_levmar(m.h, 0, 99, None, None, cfp, None, 200, 1, None, None, None) == 0
    assert L.cdm_mcmc(m.h, 0, 99, None, None, cfp, None, 10, 10, 5, 1, 0, 1, 0, None, None, None) == 0
    assert L.cdm_predicted_signal(m.h, 0, 99, 4, None, cfp, None, None, None) == 0
    assert L.cdm_bic_aic(m.h, 0, 99, 4, 0, None, None, None, cfp, None, None, None, None) == 0
    # the size checks still apply
    assert L.cdm_levmar(m.h, -1, 99, None, None, cfp, None, 200, 1, None, None, None) != 0
    assert L.cdm_mcmc(m.h, 0, 99, None, None, cfp, None, 10, 4, 5, 1, 0, 1, 0, None, None, None) != 0


def test_rician_clean_matches_reference_rule():
    """dMRI_Data.cu:14-31: non-positive entries take the smallest positive
    value when some entries are positive; all-non-positive voxels stay."""
    x = np.array([[5.0, -1.0, 0.0, 3.0, 4.0], [0.0, -2.0, 0.0, -1.0, 0.0], [1.0, 2.0, 3.0, 4.0, 5.0]],
                 dtype=np.float32)
    capi.rician_clean(x)
    np.testing.assert_array_equal(x[0], [5, 3, 3, 3, 4])
    np.testing.assert_array_equal(x[1], [0, -2, 0, -1, 0])
    np.testing.assert_array_equal(x[2], [1, 2, 3, 4, 5])


def test_tuning_table_nearest_measurement_count(tmp_path):
    """Tuned entries apply per method at the nearest measurement count;
    explicit set_group wins; a malformed table is an error."""
    t = tmp_path / "tuning.txt"
    t.write_text("# model nmeas method lanes ms\n"
                 "NODDI_Bingham 99 levmar 64 1.0\n"
                 "NODDI_Bingham 99 mcmc 4 2.0\n"
                 "NODDI_Bingham 280 mcmc 16 3.0\n"
                 "Ball_1_Stick 64 mcmc 1 1.0\n")
    m = capi.Model("NODDI_Bingham")
    assert m.load_tuning(str(t)) == 3
    assert m.get_group(capi.LEVMAR, 99) == 64
    assert m.get_group(capi.MCMC, 99) == 4
    assert m.get_group(capi.MCMC, 120) == 4
    assert m.get_group(capi.MCMC, 250) == 16
    # occupancy floor for small launches; explicit choices are kept as set
    assert m.get_group(capi.MCMC, 99, 500000) == 4
    assert m.get_group(capi.MCMC, 99, 16384) == 16
    assert m.get_group(capi.MCMC, 99, 100) == 64
    m.set_group(capi.MCMC, 64)
    assert m.get_group(capi.MCMC, 99) == 64
    m.set_group(capi.MCMC, 1)
    assert m.get_group(capi.MCMC, 99, 100) == 1
    m.set_group(capi.MCMC, 0)
    assert m.get_group(capi.MCMC, 99) == 4
    t.write_text("NODDI_Bingham 99 levmar 3\n")
    with pytest.raises(capi.CdmError, match="lanes"):
        m.load_tuning(str(t))
    t.write_text("NODDI_Bingham 99 newton 4\n")
    with pytest.raises(capi.CdmError, match="unknown method"):
        m.load_tuning(str(t))
    assert m.load_tuning(None) == 0
    t.write_text("NODDI_Bingham 99 levmar 64 3.5 2\nNODDI_Bingham 99 mcmc 4 2.0\n")
    assert m.load_tuning(str(t)) == 2
    assert m.get_occupancy(capi.LEVMAR, 99) == 2
    assert m.get_occupancy(capi.MCMC, 99) == 1
    m.set_occupancy(capi.MCMC, 4)
    assert m.get_occupancy(capi.MCMC, 99) == 4
    with pytest.raises(capi.CdmError, match="register budget"):
        m.set_occupancy(capi.MCMC, 3)
    t.write_text("NODDI_Bingham 99 levmar 64 3.5 8\n")
    with pytest.raises(capi.CdmError, match="register budget"):
        m.load_tuning(str(t))


def test_packed_triangle_index_host_build(tmp_path):
    """The packed lower-triangle index the 64-lane LM sums hand out (entry e
    to lane e % 64) maps every e back to its (i, j) for 1..32 parameters,
    NODDI_Bingham_2Fans' 14 (105 entries, two rounds of the wave) included:
    the engine header compiled for the host only and run here."""
    import subprocess
    csrc = os.path.join(os.path.dirname(capi.LIB_PATH), "..", "csrc")
    src = tmp_path / "tri.hip"
    checks = "\n".join("  ok &= check<%d>();" % n for n in range(1, 33))
    src.write_text('#include "engine/packed.h"\n#include <cstdio>\n'
                   "template <int NP> bool check() {\n"
                   "  int e = 0;\n"
                   "  for (int i = 0; i < NP; ++i)\n"
                   "    for (int j = 0; j <= i; ++j, ++e) {\n"
                   "      int a = -1, b = -1;\n"
                   "      cdm::tri_ij<NP>(e, a, b);\n"
                   "      if (a != i || b != j) { std::printf(\"NP=%d e=%d -> (%d,%d), want (%d,%d)\\n\", NP, e, a, b, i, j); return false; }\n"
                   "    }\n"
                   "  return e == cdm::tri_size(NP) && cdm::tri_ij_exact<NP>();\n"
                   "}\n"
                   "int main() {\n  bool ok = true;\n" + checks + "\n  std::printf(ok ? \"ok\\n\" : \"bad\\n\");\n"
                   "  return ok ? 0 : 1;\n}\n")
    exe = tmp_path / "tri"
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not installed")
    subprocess.run([hipcc, "--offload-host-only", "-std=c++17", "-I", csrc, str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout
