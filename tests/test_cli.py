"""Host driver (cudimot_amd/bin/cudimot): the reference's split_parts_X /
X / merge_parts_X programs.  CPU tests cover the NIfTI and part-file
formats, the voxel partition, --init_params / --FixP parts, merging and the
reference's error checks; the GPU test runs split -> fit -> merge and
compares every output with direct C-ABI calls on the same parts."""
import os
import subprocess

import numpy as np
import pytest

from tests import nifti_util as nu


def _cli():
    from cudimot_amd import build
    assert os.path.exists(build.CLI), "host driver not built (python -m cudimot_amd.build)"
    return build.CLI


def run(step, model, tmp, *args, expect=0):
    cmd = [_cli(), step, "--model=" + model] + list(args)
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
    assert r.returncode == expect, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r


def common(tmp, nparts=3, extra=()):
    return ["--data=%s/data" % tmp, "--maskfile=%s/mask" % tmp, "--partsdir=%s/parts" % tmp,
            "--outputdir=%s/out" % tmp, "--idPart=0", "--nParts=%d" % nparts, "--forcedir",
            "--logdir=%s/logs" % tmp] + list(extra)


def volume(tmp, shape=(5, 4, 3), nt=7, seed=0, gz=True, dtype=np.float32):
    rng = np.random.default_rng(seed)
    data = rng.uniform(10, 1000, size=shape + (nt,)).astype(dtype)
    mask = (rng.uniform(size=shape) > 0.3).astype(np.uint8)
    ext = ".nii.gz" if gz else ".nii"
    nu.write_nifti(str(tmp / ("data" + ext)), data)
    nu.write_nifti(str(tmp / ("mask" + ext)), mask)
    os.makedirs(tmp / "parts", exist_ok=True)
    os.makedirs(tmp / "out", exist_ok=True)
    return data.astype(np.float64), mask


@pytest.mark.parametrize("gz", [True, False])
def test_split_writes_masked_voxels_in_x_fastest_order(tmp_path, gz):
    data, mask = volume(tmp_path, gz=gz)
    run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, nparts=3))
    expect = nu.masked(data.astype(np.float32).astype(np.float64), mask)
    n = expect.shape[0]
    size = n // 3
    got = [nu.read_part(str(tmp_path / "parts" / ("part_%d" % i) / "data")) for i in range(3)]
    assert [g.shape[0] for g in got] == [size, size, n - 2 * size]  # split_parts.cc:80-94
    np.testing.assert_array_equal(np.concatenate(got), expect)
    assert os.path.exists(tmp_path / "logs" / "logfile")


def test_split_reads_scaled_int16_and_big_endian(tmp_path):
    rng = np.random.default_rng(1)
    raw = rng.integers(-300, 3000, size=(4, 3, 2, 5)).astype(np.int16)
    mask = np.ones((4, 3, 2), dtype=np.uint8)
    nu.write_nifti(str(tmp_path / "data.nii.gz"), raw, slope=0.5, inter=2.0, big_endian=True)
    nu.write_nifti(str(tmp_path / "mask.nii"), mask)
    os.makedirs(tmp_path / "parts")
    run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, nparts=1))
    got = nu.read_part(str(tmp_path / "parts/part_0/data"))
    np.testing.assert_array_equal(got, nu.masked(raw.astype(np.float64) * 0.5 + 2.0, mask))


def test_split_init_params_and_fixp_parts(tmp_path):
    data, mask = volume(tmp_path, nt=9, seed=2)
    rng = np.random.default_rng(3)
    kappa = rng.uniform(1, 10, size=mask.shape).astype(np.float32)
    s0 = rng.uniform(500, 1500, size=mask.shape).astype(np.float32)
    nu.write_nifti(str(tmp_path / "kappa.nii.gz"), kappa)
    nu.write_nifti(str(tmp_path / "S0.nii.gz"), s0)
    # NODDI_Watson: 5 parameters; only kappa given, the rest from P_init
    (tmp_path / "init.txt").write_text("\n\n%s/kappa\n\n\n" % tmp_path)
    (tmp_path / "fixp.txt").write_text("%s/S0\n" % tmp_path)
    run("split", "NODDI_Watson", tmp_path,
        *common(tmp_path, nparts=2, extra=["--init_params=%s/init.txt" % tmp_path, "--FixP=%s/fixp.txt" % tmp_path]))
    from cudimot_amd import capi
    p_init = capi.Model("NODDI_Watson").spec()["p_init"]
    for k in range(5):
        got = np.concatenate([nu.read_part(str(tmp_path / ("parts/part_%d/ParamInit_%d" % (i, k)))) for i in range(2)])
        want = nu.masked(kappa[..., None].astype(np.float64), mask) if k == 2 else np.full((got.shape[0], 1), p_init[k])
        np.testing.assert_array_equal(got, want)
    fp = np.concatenate([nu.read_part(str(tmp_path / ("parts/part_%d/FixParam_0" % i))) for i in range(2)])
    np.testing.assert_array_equal(fp, nu.masked(s0[..., None].astype(np.float64), mask))


def test_split_reference_errors(tmp_path):
    volume(tmp_path, nt=9)
    # NODDI_Watson needs --FixP (split_parts.cc:209-212)
    r = run("split", "NODDI_Watson", tmp_path, *common(tmp_path), expect=255)
    assert "--FixP" in r.stderr
    # more parts than voxels (split_parts.cc:68-71)
    r = run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, nparts=100000), expect=255)
    assert "lower than number of voxels" in r.stderr
    # no method selected (split_parts.cc:43-46)
    r = run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, extra=["--no_LevMar"]), expect=255)
    assert "at least one method" in r.stderr
    # too many --init_params lines
    (tmp_path / "init.txt").write_text("\n" * 7)
    r = run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, extra=["--init_params=%s/init.txt" % tmp_path]),
            expect=255)
    assert "Too many lines" in r.stderr
    # a compulsory option missing: usage and status 2 (cudimotoptions.cc:31-35)
    r = run("split", "Ball_1_Stick", tmp_path, "--data=x", expect=2)
    assert "missing compulsory option --maskfile" in r.stderr
    r = run("split", "Ball_1_Stick", tmp_path, *common(tmp_path, extra=["--no_such_option"]), expect=2)


@pytest.mark.parametrize("mcmc", [False, True])
def test_merge_joins_parts_into_masked_volumes(tmp_path, mcmc):
    _, mask = volume(tmp_path)
    nvox = int((mask > 0).sum())
    P = 5  # Ball_1_Stick
    ns = 4 if mcmc else 1
    counts = [nvox // 3, nvox // 3, nvox - 2 * (nvox // 3)]
    rng = np.random.default_rng(5)
    full = {}
    names = ["Param_%d_samples" % k for k in range(P)] + ["PredictedSignal", "BIC", "AIC"]
    rows = {n: (7 if n == "PredictedSignal" else 1 if n in ("BIC", "AIC") else ns) for n in names}
    if mcmc:
        names.append("Tau_samples")
        rows["Tau_samples"] = ns
    for n in names:
        full[n] = rng.normal(size=(nvox, rows[n])).astype(np.float32).astype(np.float64)
    a = 0
    for i, c in enumerate(counts):
        os.makedirs(tmp_path / ("parts/part_%d" % i), exist_ok=True)
        for n in names:
            nu.write_part(str(tmp_path / ("parts/part_%d/%s" % (i, n))), full[n][a:a + c])
        a += c
    extra = ["--getPredictedSignal", "--BIC_AIC"] + (["--runMCMC", "--nj=8", "--se=2", "--rician"] if mcmc else [])
    run("merge", "Ball_1_Stick", tmp_path, *common(tmp_path, extra=extra))
    assert not os.path.exists(tmp_path / "parts")  # removed without --keepTmp (merge_parts.cc:142-146)
    for n in names:
        x, h = nu.read_nifti(str(tmp_path / "out" / (n + ".nii.gz")))
        assert x.shape == mask.shape + (rows[n],)
        np.testing.assert_array_equal(nu.masked(x, mask), full[n])
        assert np.all(x[mask == 0] == 0)
        assert h["datatype"] == 16 and h["magic"] == b"n+1\0"
        assert h["cal_max"] == np.float32(x.max()) and h["cal_min"] == np.float32(x.min())  # merge_parts.cc:57-59
        assert h["pixdim"][1:4] == (2.0, 2.0, 2.0)  # the mask's geometry


def test_merge_keeps_parts_and_row_major_order(tmp_path):
    _, mask = volume(tmp_path)
    nvox = int((mask > 0).sum())
    vals = np.arange(nvox * 5 * 3, dtype=np.float64).reshape(nvox, 5, 3)
    os.makedirs(tmp_path / "parts/part_0")
    for k in range(5):
        nu.write_part(str(tmp_path / ("parts/part_0/Param_%d_samples" % k)), vals[:, k, :], row_major=True)
    raw = open(tmp_path / "parts/part_0/Param_0_samples", "rb").read()
    assert np.frombuffer(raw, dtype="<f8", offset=16)[1] == vals[1, 0, 0]  # row-major: voxels of sample 0 first
    run("merge", "Ball_1_Stick", tmp_path,
        *common(tmp_path, nparts=1, extra=["--keepTmp", "--part_rowmajor", "--runMCMC", "--nj=3", "--se=1"]))
    assert os.path.exists(tmp_path / "parts/part_0")
    for k in range(5):
        x, _ = nu.read_nifti(str(tmp_path / "out" / ("Param_%d_samples.nii.gz" % k)))
        np.testing.assert_array_equal(nu.masked(x, mask), vals[:, k, :])


def test_merge_rejects_wrong_sample_count(tmp_path):
    _, mask = volume(tmp_path)
    nvox = int((mask > 0).sum())
    os.makedirs(tmp_path / "parts/part_0")
    for k in range(5):
        nu.write_part(str(tmp_path / ("parts/part_0/Param_%d_samples" % k)), np.zeros((nvox, 3)))
    r = run("merge", "Ball_1_Stick", tmp_path, *common(tmp_path, nparts=1), expect=255)
    assert "not correct" in r.stderr


def _acq_files(tmp, cfp):
    np.savetxt(str(tmp / "bvecs"), cfp[:, :3].T, fmt="%.17g")
    np.savetxt(str(tmp / "bvals"), cfp[:, 3:4].T, fmt="%.17g")
    (tmp / "cfp.txt").write_text("%s/bvecs\n%s/bvals\n" % (tmp, tmp))


@pytest.mark.gpu
@pytest.mark.parametrize("mcmc", [False, True])
def test_split_fit_merge_equals_direct_calls(tmp_path, mcmc):
    import torch
    from cudimot_amd import capi, synth
    from tests.helpers import dev, host
    model, nparts = "Ball_1_Stick", 2
    cfp = synth.acquisition(4, [(1000, 30), (2000, 30)], seed=21).astype(np.float32).astype(np.float64)
    shape = (12, 10, 6)
    rng = np.random.default_rng(22)
    mask = (rng.uniform(size=shape) > 0.25).astype(np.uint8)
    nvox = int(mask.sum())
    t = synth.truth_ball_1_stick(nvox, seed=23)
    o = capi.Model(model)
    clean = torch.empty((nvox, cfp.shape[0]), device="cuda")
    o.predicted_signal(dev(t.reshape(nvox, -1, 1)), dev(cfp), None, clean)
    c = host(clean)
    sig = (t[:, :1] / 30.0)
    noisy = np.sqrt((c + sig * rng.normal(size=c.shape)) ** 2 + (sig * rng.normal(size=c.shape)) ** 2)
    # in-mask voxels in x-fastest order
    flat = np.zeros((int(np.prod(shape)), cfp.shape[0]), dtype=np.float32)
    flat[mask.reshape(-1, order="F") > 0] = noisy
    vol = flat.reshape(shape + (cfp.shape[0],), order="F")
    nu.write_nifti(str(tmp_path / "data.nii.gz"), vol)
    nu.write_nifti(str(tmp_path / "mask.nii.gz"), mask)
    _acq_files(tmp_path, cfp)
    os.makedirs(tmp_path / "parts")
    os.makedirs(tmp_path / "out")
    opts = ["--CFP=%s/cfp.txt" % tmp_path, "--getPredictedSignal", "--BIC_AIC", "--iterLevMar=50"]
    if mcmc:
        opts += ["--runMCMC", "--bi=40", "--nj=30", "--se=10", "--seed=77", "--debug=5"]
    run("split", model, tmp_path, *common(tmp_path, nparts, opts))
    for i in range(nparts):
        a = common(tmp_path, nparts, opts)
        a[4] = "--idPart=%d" % i
        print("fit part", i, flush=True)
        r = run("fit", model, tmp_path, *a)
        if mcmc:  # --debug=5: voxel 5 of each part, and the fit itself unchanged (below)
            assert "----- Levenberg_Marquardt GPU algorithm: voxel 5 -----" in r.stdout, r.stdout[-2000:]
            assert "----- MCMC GPU algorithm: voxel 5 -----" in r.stdout
            assert r.stdout.count("Jump ") == 70
    run("merge", model, tmp_path, *common(tmp_path, nparts, opts))

    # the same parts through the C-ABI
    meas_all = nu.masked(vol.astype(np.float64), mask).astype(np.float32)
    size = nvox // nparts
    p_init = np.asarray(o.spec()["p_init"], dtype=np.float32)
    ns = 3 if mcmc else 1
    want = {k: [] for k in ["samples", "pred", "bic", "aic"]}
    for i in range(nparts):
        a = i * size
        b = nvox if i == nparts - 1 else a + size
        n = b - a
        m = capi.Model(model)
        meas = dev(meas_all[a:b])
        p = dev(np.tile(p_init, (n, 1)))
        d_cfp = dev(cfp.astype(np.float32))
        m.levmar(meas, p, d_cfp, None, max_iters=50)
        if mcmc:
            s = torch.empty((n, 5, ns), device="cuda")
            m.mcmc(meas, p, d_cfp, None, s, None, nburn=40, njumps=30, sampleevery=10, seed=77, voxel_offset=a)
        else:
            s = p.reshape(n, 5, 1)
        pred = torch.empty((n, cfp.shape[0]), device="cuda")
        m.predicted_signal(s, d_cfp, None, pred)
        bic, aic = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
        m.bic_aic(meas, s, d_cfp, None, bic, aic)
        capi.sync()
        want["samples"].append(host(s))
        want["pred"].append(host(pred))
        want["bic"].append(host(bic))
        want["aic"].append(host(aic))
    samples = np.concatenate(want["samples"])
    for k in range(5):
        x, _ = nu.read_nifti(str(tmp_path / "out" / ("Param_%d_samples.nii.gz" % k)))
        np.testing.assert_array_equal(nu.masked(x, mask), samples[:, k, :])
    x, _ = nu.read_nifti(str(tmp_path / "out/PredictedSignal.nii.gz"))
    np.testing.assert_array_equal(nu.masked(x, mask), np.concatenate(want["pred"]))
    for name, key in (("BIC", "bic"), ("AIC", "aic")):
        x, _ = nu.read_nifti(str(tmp_path / "out" / (name + ".nii.gz")))
        np.testing.assert_array_equal(nu.masked(x, mask)[:, 0], np.concatenate(want[key]))


def test_finish_usage_errors(tmp_path):
    r = subprocess.run([_cli(), "finish", "--model=NODDI_Bingham"], capture_output=True, text=True)
    assert r.returncode == 2 and "output directory" in r.stderr
    r = subprocess.run([_cli(), "finish", "--model=toymodel", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 255 and "no finish step" in r.stderr


@pytest.mark.gpu
def test_finish_noddi_bingham_maps_match_oracle(tmp_path):
    from oracle import finish_ref as fr
    from tests.test_finish import _stick_samples
    shape = (6, 5, 4)
    rng = np.random.default_rng(31)
    mask = (rng.uniform(size=shape) > 0.3).astype(np.uint8)
    nvox, S = int(mask.sum()), 20
    x = _stick_samples(nvox, S, 32)  # [nvox][7][S]: fiso, fintra, kappa, beta, th, ph, psi
    nu.write_nifti(str(tmp_path / "nodif_brain_mask.nii.gz"), mask)
    sel = mask.reshape(-1, order="F") > 0
    for k in range(7):
        flat = np.zeros((int(np.prod(shape)), S), dtype=np.float32)
        flat[sel] = x[:, k, :]
        nu.write_nifti(str(tmp_path / ("Param_%d_samples.nii.gz" % k)), flat.reshape(shape + (S,), order="F"))
    r = subprocess.run([_cli(), "finish", "--model=NODDI_Bingham", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def get(name):
        v, _ = nu.read_nifti(str(tmp_path / (name + ".nii.gz")))
        return nu.masked(v, mask)
    names = ["fiso", "fintra", "kappa", "beta", "th", "ph", "psi"]
    for k, n in enumerate(names):  # renamed, unchanged
        np.testing.assert_array_equal(get(n + "_samples"), x[:, k, :])
        assert not os.path.exists(tmp_path / ("Param_%d_samples.nii.gz" % k))
    mean = fr.sample_mean(x)
    for k in range(4):
        np.testing.assert_allclose(get("mean_" + names[k])[:, 0], mean[:, k], rtol=1e-6)
    np.testing.assert_array_equal(get("k2_samples"), (x[:, 2, :] - x[:, 3, :]).astype(np.float32))
    maps = fr.noddi_maps(x, 2, 3)
    for i, n in enumerate(["mean_k2", "ODIp", "ODIs", "ODItot", "DA"]):
        np.testing.assert_allclose(get(n)[:, 0], maps[:, i], rtol=2e-6, atol=1e-7)
    for name, psi in (("dyads1", None), ("fanningDir", 6)):
        d, disp = fr.dyads(x, 4, 5, psi)
        np.testing.assert_allclose(np.abs(np.sum(get(name) * d, axis=1)), 1.0, atol=2e-5)
        if psi is None:
            np.testing.assert_allclose(get("dyads1_dispersion")[:, 0], disp, atol=2e-6)
