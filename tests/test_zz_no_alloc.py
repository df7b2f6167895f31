"""om_flux_dev_noalloc: the allocation-free variant of om_flux_dev.

It runs the same kernels in the same order on the library's static device
scratch, so its values, levels, convergence flags and E-fields must equal
om_flux_dev's bit for bit (the bar for this test: identical).  The library
throws if any device or pinned allocation is attempted while it runs, so a
fresh context whose only calls after the trajectory are allocation-free proves
that none of its paths allocate.  Two calls in a row, the second smaller; a
request beyond the scratch fails cleanly; accumulate is refused."""
import numpy as np
import pytest

from tests.common import planar

gpu = pytest.mark.gpu


def _grid(n, w=0.01):
    from oscars_amd import geometry as G
    obs, _, _ = G.rectangle("XY", n, n, w, w, [0, 0, 0], [0, 0, 30], 0, 2)
    return np.array(obs)


def _call(dev, fn, obs, e, mode):
    import torch
    obs_t = torch.tensor(obs, dtype=torch.float64, device="cuda")
    e_t = torch.tensor(e, dtype=torch.float64, device="cuda")
    n = len(obs) * len(e)
    out = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    lev = torch.full((n,), -9, dtype=torch.int32, device="cuda")
    cv = torch.full((n,), 9, dtype=torch.uint8, device="cuda")
    ef = torch.zeros(n, 6, dtype=torch.float64, device="cuda")
    st = fn(0, obs_t, e_t, [0.0] * 6, dev.conv(mode=mode), out, lev, cv, ef)
    return out.cpu().numpy(), lev.cpu().numpy(), cv.cpu().numpy(), ef.cpu().numpy(), st


@gpu
@pytest.mark.parametrize("mode", ["plain", "tiled", "mixed", "point_mixed"])
def test_noalloc_matches_flux_dev(mode):
    from oscars_amd import _lib
    from oscars_amd.sr import sr
    s = sr()
    planar(s, npoints=8001)
    s.calculate_trajectory()
    ref_dev = s._device()
    # a second context holding the same trajectory: after om_trajectories it
    # only ever calls the allocation-free entry point
    fresh = _lib.Device(0)
    fresh.trajectories(s._fields, [s._om_particle(s._particle)])
    cases = [(_grid(21), np.linspace(120.0, 900.0, 9)), (_grid(7), np.array([152.11, 455.0]))]
    for k, (obs, e) in enumerate(cases):
        a = _call(ref_dev, ref_dev.flux_dev, obs, e, mode)
        b = _call(fresh, fresh.flux_dev_noalloc, obs, e, mode)
        print("%s call %d: %d points, evals %g / %g, rechecked %d / %d" %
              (mode, k, len(a[0]), a[4]["evals"], b[4]["evals"], a[4]["n_rechecked"], b[4]["n_rechecked"]))
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y), (mode, k)
        assert a[4]["evals"] == b[4]["evals"] and a[4]["max_level"] == b[4]["max_level"]
        assert a[4]["n_unconverged"] == b[4]["n_unconverged"] and a[4]["n_rechecked"] == b[4]["n_rechecked"]


@gpu
def test_noalloc_limits():
    import torch
    from oscars_amd import _lib
    from oscars_amd.sr import sr
    s = sr()
    planar(s, npoints=4001)
    dev = _lib.Device(0)
    dev.trajectories(s._fields, [s._om_particle(s._particle)])
    obs = _grid(5)
    obs_t = torch.tensor(obs, dtype=torch.float64, device="cuda")
    e_t = torch.tensor([200.0, 300.0], dtype=torch.float64, device="cuda")
    out = torch.zeros(len(obs) * 2, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception, match="accumulate is not supported"):
        dev.flux_dev_noalloc(0, obs_t, e_t, [0.0] * 6, dev.conv(mode="tiled", accumulate=True), out)
    # 2^20 points x 8 energies needs far more than the 64 MiB scratch
    big = torch.zeros(1 << 20, 3, dtype=torch.float64, device="cuda")
    big[:, 2] = 30.0
    e8 = torch.linspace(100.0, 800.0, 8, dtype=torch.float64, device="cuda")
    out8 = torch.zeros((1 << 20) * 8, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception, match="needs more device scratch"):
        dev.flux_dev_noalloc(0, big, e8, [0.0] * 6, dev.conv(mode="tiled"), out8)
    # the context still works afterwards
    st = dev.flux_dev_noalloc(0, obs_t, e_t, [0.0] * 6, dev.conv(mode="tiled"), out)
    ref = torch.zeros_like(out)
    dev.flux_dev(0, obs_t, e_t, [0.0] * 6, dev.conv(mode="tiled"), ref)
    assert torch.equal(out, ref) and st["evals"] > 0
