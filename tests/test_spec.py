"""modelpriors / grid / --fixed parsing through the C-ABI, with the
reference's rules (Model.cu:16-393)."""
import os

import pytest

from cudimot_amd import capi

REF = "/root/reference/mymodels"


def parse(text, model="Ball_2_Sticks", tmp_path=None):
    p = tmp_path / "priors"
    p.write_text(text)
    return capi.Model(model, str(p)).spec()


def test_shipped_priors_match_reference_values():
    s = capi.Model("Ball_1_Stick").spec()
    assert s["p_init"] == [1, 0.001, 0.6, 1.5, 1.5]
    assert s["bound_types"] == [1, 3, 3, 3, 3]  # (1e-9,) is min-only
    assert s["bmin"][:3] == [1e-9, 1e-9, 0.01] and s["bmax"][1:3] == [0.01, 1]
    assert s["prior_types"] == [0, 0, 0, 4, 0]
    s = capi.Model("Ball_2_Sticks").spec()
    assert s["prior_types"] == [0, 0, 0, 4, 0, 3, 4, 0] and s["pa"][5] == 1.0
    s = capi.Model("DiffusionTensor").spec()
    assert s["bound_types"] == [3, 3, 0, 0, 3, 0, 3]
    s = capi.Model("NODDI_Bingham").spec()
    # "priors[4]=sin()" is not recognised (quirk kept): no prior on th
    assert s["prior_types"] == [0] * 7
    assert s["bmin"][2:4] == [0.1, 0.01] and s["bmax"][2:4] == [64, 32]
    s = capi.Model("toymodel").spec()
    assert s["init_provided"] and s["bound_types"] == [0, 0]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout absent")
@pytest.mark.parametrize("model", ["Ball_1_Stick", "Ball_2_Sticks", "DiffusionTensor", "NODDI_Watson",
                                   "NODDI_Bingham", "NODDI_Bingham_2Fans", "toymodel"])
def test_reference_modelpriors_parse_like_ours(model):
    ours = capi.Model(model).spec()
    ref = capi.Model(model, os.path.join(REF, model, "modelpriors")).spec()
    assert ours == ref


def test_comment_marker_anywhere_skips_line(tmp_path):
    s = parse("P_init[1,2,3,4,5,6,7,8]\nbounds[1]=(0,5) // trailing comment\nbounds[2]=(0,1)\n", tmp_path=tmp_path)
    assert s["bound_types"][1] == 0 and s["bound_types"][2] == 3


def test_bound_forms(tmp_path):
    s = parse("bounds[0]=(2,)\nbounds[1]=(,7)\nbounds[2]=(-1,1)\n", tmp_path=tmp_path)
    assert s["bound_types"][:3] == [1, 2, 3]
    assert s["bmin"][0] == 2 and s["bmax"][1] == 7 and (s["bmin"][2], s["bmax"][2]) == (-1, 1)


def test_prior_forms_and_priors_spelling(tmp_path):
    s = parse("prior[0]=Gaussian(3,0.5)\nprior[1]=Gamma(2,4)\nprior[2]=ARD(0.5)\nprior[3]=sin()\n"
              "prior[4]=custom()\npriors[5]=sin()\n", tmp_path=tmp_path)
    assert s["prior_types"][:6] == [1, 2, 3, 4, 5, 0]
    assert (s["pa"][0], s["pb"][0]) == (3, 0.5) and (s["pa"][1], s["pb"][1]) == (2, 4) and s["pa"][2] == 0.5


@pytest.mark.parametrize("text,msg", [
    ("P_init[1,2,3]\n", "do not match"),
    ("bounds[9]=(0,1)\n", "Wrong number of parameter"),
    ("prior[1]=Cauchy(1,2)\n", "not a recognised prior"),
    ("prior[1]=Gaussian(,2)\n", "needs a first argument"),
    ("bounds[1]=(,)\n", "at least a min or a max"),
])
def test_priors_errors(tmp_path, text, msg):
    p = tmp_path / "priors"
    p.write_text(text)
    with pytest.raises(capi.CdmError, match=msg):
        capi.Model("Ball_2_Sticks", str(p))


def test_missing_priors_file():
    with pytest.raises(capi.CdmError, match="Unable to open"):
        capi.Model("Ball_1_Stick", "/nonexistent/priors")


def test_grid_cartesian_order_lowest_parameter_outermost():
    m = capi.Model("Ball_2_Sticks")
    gp, g = m.load_grid("search[5]=(0.1,0.2)\n// search[3]=(9)\nsearch[2]=(0.5,0.6,0.7)\nsearch[2]=(0.8)\n",
                        is_text=True)
    assert gp == [2, 5]
    assert g == [[0.5, 0.1], [0.5, 0.2], [0.6, 0.1], [0.6, 0.2], [0.7, 0.1], [0.7, 0.2], [0.8, 0.1], [0.8, 0.2]]


def test_grid_reference_pipeline_sizes():
    m = capi.Model("NODDI_Bingham")
    gp, g = m.load_grid("search[0]=(0.0,0.05,0.1,0.2,0.3)\nsearch[1]=(0.3,0.5,0.6,0.7,0.75,0.8,0.85,0.9,0.95,1.0)\n"
                        "search[2]=(0.5,1,2,3,4,5,6,7,8,10,12,15,20,25,30,40,50)\n"
                        "search[3]=(0.2,0.5,1.0,1.5,2.0,2.5,3.0,4.5,6.5,9,12,16,20)\n", is_text=True)
    assert gp == [0, 1, 2, 3] and len(g) == 11050
    assert g[0] == [0.0, 0.3, 0.5, 0.2] and g[-1] == [0.3, 1.0, 50, 20]


def test_fixed_list():
    m = capi.Model("NODDI_Bingham")
    m.set_fixed("4,5")
    assert m.spec()["fixed"] == [0, 0, 0, 0, 1, 1, 0]
    with pytest.raises(capi.CdmError, match="Wrong number of parameter 9"):
        m.set_fixed("1,9")
