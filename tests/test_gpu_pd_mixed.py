"""Mixed power density (OM_MODE_MIXED): reciprocal-based terms, with the
points whose stop decision lies within the terms' error bound recomputed in
FP64.  Checked against the FP64 GPU path on whole grids (levels identical,
values within 1e-5 relative, denominator floored at 1e-6 of the map maximum)
and against the oracle on sampled points of the 401 x 401 map (1e-5)."""
import numpy as np
import pytest

from tests.common import epu, rel_err

gpu = pytest.mark.gpu


@pytest.fixture(scope="module")
def epu_dev():
    from oscars_amd.sr import sr
    g = sr()
    epu(g)
    g.calculate_trajectory()
    return g, g._device()


def _rect(n, w, z):
    from oscars_amd import geometry as G
    obs, _, _ = G.rectangle("XY", n, n, w, w, [0, 0, 0], [0, 0, z], 0, 2)
    return np.array(obs)


@gpu
def test_mixed_power_density_401_matches_fp64_and_oracle(epu_dev):
    from oracle.pyoracle import Oracle
    g, dev = epu_dev
    obs = _rect(401, 0.08, 20.0)
    v64, l64, c64, s64 = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(mode="plain"))
    vm, lm, cm, sm = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(mode="mixed"))
    e = rel_err(vm, v64, 1e-6)
    mism = int(np.sum(lm != l64))
    print("401x401 power density: mixed vs FP64 max rel err %.3e, level mismatches %d, rechecked %d, "
          "kernel ms FP64 %.3f mixed %.3f" % (e, mism, sm["n_rechecked"], s64["kernel_ms"], sm["kernel_ms"]))
    assert mism == 0
    assert np.array_equal(cm, c64)
    assert e < 1e-5
    o = Oracle(nthreads=8)
    epu(o)
    o.calculate_trajectory()
    rng = np.random.default_rng(3)
    idx = np.unique(np.concatenate([[0, len(obs) // 2, len(obs) - 1], rng.choice(len(obs), 97, replace=False)]))
    ov, olev, _ = o.power_density_points(obs[idx], [0, 0, 1], directional=False, carry_quirk=False)
    eo = float(np.max(np.abs(vm[idx] - ov) / np.maximum(np.abs(ov), 1e-6 * np.max(np.abs(v64)))))
    print("mixed vs oracle on %d points: %.3e" % (len(idx), eo))
    assert np.array_equal(lm[idx], olev)
    assert eo < 1e-5


@gpu
@pytest.mark.parametrize("surface", ["directional", "normals", "line"])
def test_mixed_power_density_surfaces(epu_dev, surface):
    g, dev = epu_dev
    rng = np.random.default_rng(8)
    obs = _rect(41, 0.02, 20.0)
    if surface == "directional":
        run = lambda mode: dev.power_density(0, obs, [0, 0, 1], 1, dev.conv(mode=mode))
    else:
        if surface == "normals":
            nv = rng.normal(size=(len(obs), 3))
            nv[:, 2] = np.abs(nv[:, 2]) + 0.3
            nv /= np.linalg.norm(nv, axis=1)[:, None]
            nv[::7, 2] *= -1  # some facing away
        else:
            nv = None
        run = lambda mode: dev.power_density_surface(0, obs, nv, 1 if surface == "normals" else 0,
                                                      dev.conv(mode=mode))
    v64, l64, _, _ = run("plain")
    vm, lm, _, sm = run("mixed")
    e = rel_err(vm, v64, 1e-6)
    print("%s: mixed vs FP64 %.3e, rechecked %d" % (surface, e, sm["n_rechecked"]))
    assert np.array_equal(lm, l64)
    assert e < 1e-5


@gpu
def test_mixed_power_density_rechecks_a_borderline_decision(epu_dev):
    """precision set to one point's final relative change (from the FP64
    path): that point's decision is borderline in the mixed path, is
    recomputed in FP64, and stops where the FP64 path stops."""
    g, dev = epu_dev
    obs = _rect(21, 0.02, 20.0)
    pr, _, _, _ = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(quantity=1, mode="plain"))
    k = int(np.argmax(pr))
    prec = float(pr[k])
    v64, l64, c64, _ = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(precision=prec, mode="plain"))
    vm, lm, cm, sm = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(precision=prec, mode="mixed"))
    print("precision %.17g: rechecked %d, level of point %d: %d" % (prec, sm["n_rechecked"], k, l64[k]))
    assert sm["n_rechecked"] >= 1
    assert np.array_equal(lm, l64)
    assert np.array_equal(cm, c64)
    assert rel_err(vm, v64, 1e-6) < 1e-5


@gpu
def test_mixed_power_density_capped(epu_dev):
    """max_level below convergence: unconverged points, flags and levels as
    the FP64 path's; values within 1e-5, the precision output (a relative
    change, ~1e-3) within 1e-6 absolute."""
    g, dev = epu_dev
    obs = _rect(9, 0.08, 20.0)
    for q in (0, 1, 2):
        v64, l64, c64, _ = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(max_level=9, quantity=q, mode="plain"))
        vm, lm, cm, _ = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(max_level=9, quantity=q, mode="mixed"))
        assert not c64.all()
        assert np.array_equal(lm, l64) and np.array_equal(cm, c64)
        if q == 1:  # the last relative change, a difference of two sums: absolute bar
            assert np.max(np.abs(vm - v64)) < 1e-6
        else:
            assert rel_err(vm, v64, 1e-6) < 1e-5


@gpu
def test_mixed_power_density_device_io(epu_dev):
    """om_power_density_dev (device tensors) gives the host entry point's
    values, levels and flags bit for bit, in both modes."""
    import torch
    g, dev = epu_dev
    obs = _rect(31, 0.04, 20.0)
    t_obs = torch.tensor(obs, dtype=torch.float64, device="cuda")
    for mode in ("plain", "mixed"):
        v, lev, cv, st = dev.power_density(0, obs, [0, 0, 1], 0, dev.conv(mode=mode))
        out = torch.zeros(len(obs), dtype=torch.float64, device="cuda")
        lt = torch.zeros(len(obs), dtype=torch.int32, device="cuda")
        ct = torch.zeros(len(obs), dtype=torch.uint8, device="cuda")
        st2 = dev.power_density_dev(0, t_obs, [0, 0, 1], 0, dev.conv(mode=mode), out, lt, ct)
        assert np.array_equal(out.cpu().numpy(), v)
        assert np.array_equal(lt.cpu().numpy(), lev) and np.array_equal(ct.cpu().numpy(), cv)
        assert st2["n_rechecked"] == st["n_rechecked"]
