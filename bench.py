"""Benchmark: Lienard-Wiechert integrand evaluations per second on one MI355X.

Workload (one step), a planar undulator (By 1.0 T, period 49 mm, 31 periods,
3 GeV, 0.5 A), trajectory z in [-1, 1] m with 20001 RK4 points:
  * trajectory: RK4 + spline on the device, level samples rebuilt every step
  * flux map 201 x 201 on +-10 mm at 30 m, at 152.11 eV (first harmonic) and
    at 260 eV (the literal energy in the configuration)
  * flux sweep 51 x 51 on the same rectangle x 512 energies, 100-3000 eV
One "eval" is one (observation point, energy, trajectory sample) integrand
evaluation; a point converged at level L costs 2^(L+1)-1 evals.

The headline mode (default "mixed") runs the single-energy maps with the FP64
per-point kernel (with one energy there is nothing to share across lanes, the
amplitude dominates, and the FP32 sin/cos variant is no faster once its
borderline points are rechecked) and the sweep with the observation x energy
tiled kernel with FP32 sin/cos of the exact FP64 phase.  The plain FP64 path (one lane per output
point, everything FP64, the reference-faithful baseline) is timed on the same
workload for the ratio, and the mixed result is compared with it point by point.

Prints ONE JSON line (rank 0, N=1).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP64 flops per eval (divisions and square roots count as one flop, sin/cos counted apart):
#   per-point kernel: amplitude + phase 76, phase-step bookkeeping 3
#   tiled kernel: phase 1, phase-step bookkeeping 2, complex MAC 12 per eval, plus 66 per (obs, sample)
FLOPS_POINT = 79.0
FLOPS_TILE_EVAL = 15.0
FLOPS_TILE_SAMPLE = 66.0
PEAK_FP64_TFLOPS = 78.6


def scenario(s):
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(20001)


def rect(n, w=0.02):
    from oscars_amd import geometry as G
    obs, nv, cont = G.rectangle("XY", n, n, w, w, [0, 0, 0], [0, 0, 30], 0, 2)
    return np.array(obs)


class Workload:
    """The step's inputs and outputs live in HBM (torch tensors) before the timed
    region: the flux calls take device pointers (om_flux_dev), so a step moves
    no grid or result over PCIe."""

    def __init__(self, sr_obj, n_map=201, n_sweep=51, n_e=512):
        import torch
        self.s = sr_obj
        self.obs_map = rect(n_map)
        self.obs_sweep = rect(n_sweep)
        self.e_sweep = np.array([100 + (3000 - 100) / (n_e - 1) * i for i in range(n_e)])
        self.map_energies = [152.11, 260.0]
        self.only_sweep = False
        cuda = dict(device="cuda")
        self.t_obs_map = torch.tensor(self.obs_map, dtype=torch.float64, **cuda)
        self.t_obs_sweep = torch.tensor(self.obs_sweep, dtype=torch.float64, **cuda)
        self.t_e_sweep = torch.tensor(self.e_sweep, dtype=torch.float64, **cuda)
        self.t_e_map = {e: torch.tensor([e], dtype=torch.float64, **cuda) for e in self.map_energies}
        self.outs = {}
        for k, n in [("map_%g" % e, len(self.obs_map)) for e in self.map_energies] + \
                    [("sweep", len(self.obs_sweep) * n_e)]:
            self.outs[k] = (torch.zeros(n, dtype=torch.float64, **cuda), torch.zeros(n, dtype=torch.int32, **cuda),
                            torch.zeros(n, dtype=torch.uint8, **cuda))

    def traj(self):
        """RK4 + spline on the device (returns once queued on this context's stream)."""
        s = self.s
        s._device().trajectories(s._fields, [s._om_particle(s._particle)], check_status=False)

    def step(self, mode):
        """One pass; returns dict of per-phase evals / kernel ms / output tensors (device)."""
        self.traj()
        return self.sums(mode)

    def sums(self, mode):
        """The step's flux sums on the resident trajectory (level samples are rebuilt here)."""
        s = self.s
        dev = s._device()
        pol = [0.0] * 6
        out = {"evals": 0.0, "kernel_ms": {}, "values": {}, "levels": {}, "rechecked": {}}
        map_mode = "plain"  # see the module docstring
        for e in ([] if self.only_sweep else self.map_energies):
            k = "map_%g" % e
            v, lev, cv = self.outs[k]
            st = dev.flux_dev(0, self.t_obs_map, self.t_e_map[e], pol, dev.conv(mode=map_mode), v, lev, cv)
            out["evals"] += st["evals"]
            out["kernel_ms"][k] = st["kernel_ms"]
            out["values"][k] = v
            out["levels"][k] = lev
            out["rechecked"][k] = st["n_rechecked"]
        sweep_mode = {"plain": "plain", "tiled": "tiled", "mixed": "mixed"}[mode]
        v, lev, cv = self.outs["sweep"]
        st = dev.flux_dev(0, self.t_obs_sweep, self.t_e_sweep, pol, dev.conv(mode=sweep_mode), v, lev, cv)
        out["evals"] += st["evals"]
        out["sweep_evals"] = st["evals"]
        out["kernel_ms"]["sweep"] = st["kernel_ms"]
        out["values"]["sweep"] = v
        out["levels"]["sweep"] = lev
        out["rechecked"]["sweep"] = st["n_rechecked"]
        return out


def host_api_step_ms(work, mode, steps=2):
    """ms per step of the same workload through the host-buffer entry points
    (om_flux: grids uploaded and results copied back over PCIe every call)."""
    import torch
    s = work.s
    dev = s._device()
    pol = [0.0] * 6
    map_mode = "plain"
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.trajectories(s._fields, [s._om_particle(s._particle)])
        for e in work.map_energies:
            dev.flux(0, work.obs_map, [e], pol, dev.conv(mode=map_mode))
        dev.flux(0, work.obs_sweep, work.e_sweep, pol, dev.conv(mode=mode))
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps


def to_host(res):
    """Copies a step's device outputs to numpy (outside the timed region)."""
    for key in ("values", "levels"):
        res[key] = {k: t.cpu().numpy() for k, t in res[key].items()}
    res["sweep_levels"] = res["levels"]["sweep"]
    return res


def run_steps(works, mode, n):
    """n steps.  With two workloads (two contexts, each on its own HIP stream) the
    steps are software-pipelined: the trajectory of step k+1 is queued on the
    other context before step k's sums, so the serial RK4 (one wave) runs beside
    the sums instead of before them.  Every step still integrates its own
    trajectory and computes all its sums."""
    res, evals, kms = None, 0.0, {}
    if n <= 0:
        return res, evals, kms
    if len(works) == 2:
        works[0].traj()
        for k in range(n):
            if k + 1 < n:
                works[(k + 1) % 2].traj()
            res = works[k % 2].sums(mode)
            evals += res["evals"]
            for key, v in res["kernel_ms"].items():
                kms[key] = kms.get(key, 0.0) + v
    else:
        for _ in range(n):
            res = works[0].step(mode)
            evals += res["evals"]
            for key, v in res["kernel_ms"].items():
                kms[key] = kms.get(key, 0.0) + v
    return res, evals, kms


def timed(works, mode, steps, warmup):
    import torch
    run_steps(works, mode, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, evals, kms = run_steps(works, mode, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0), evals, {k: v / steps for k, v in kms.items()}, to_host(res)


def rel_errors(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b)
    pointwise = np.abs(a - b) / np.where(den > 0, den, 1.0)
    floored = np.abs(a - b) / np.maximum(den, 1e-3 * np.max(den))
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    return dict(max_pointwise=float(np.max(pointwise)), max_floored_1e3=float(np.max(floored)), l2=float(l2))


def cpu_baseline(work, sweep_levels, n_e_sample=8):
    """Reference C++ core (oracle/_ref/libref.so, built from the reference
    sources) on this host, nthreads = host cores, on a bounded sample: the
    201x201 map at 152.11 eV and the 51x51 sweep at n_e_sample energies.
    Falls back to the C restatement (kind 'port') when the reference build is absent."""
    from oracle import pyoracle as P
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    kind = "reference" if P.ref_available() else "port"
    sim = P.RefSim(nthreads=ncores) if kind == "reference" else P.Oracle(nthreads=ncores)
    scenario(sim)
    if kind == "port":
        sim.calculate_trajectory()
    idx = np.linspace(0, len(work.e_sweep) - 1, n_e_sample).round().astype(int)
    levels = sweep_levels.reshape(len(work.obs_sweep), len(work.e_sweep))[:, idx]
    t0 = time.perf_counter()
    if kind == "reference":
        _, _, _ = sim.calculate_flux_rectangle(152.11, [201, 201], [0.02, 0.02], translation=[0, 0, 30])
    else:
        sim.calculate_flux_rectangle(152.11, [201, 201], [0.02, 0.02], translation=[0, 0, 30])
    t_map = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k, e in enumerate(work.e_sweep[idx]):
        sim.calculate_flux_rectangle(float(e), [51, 51], [0.02, 0.02], translation=[0, 0, 30])
    t_sweep = time.perf_counter() - t0
    return dict(t_map=t_map, t_sweep=t_sweep, sweep_evals=float(np.sum(2.0 ** (levels + 1) - 1)), kind=kind,
                cores=ncores, n_e=n_e_sample)


def batched_bench(args):
    """Config: 2048 electrons from a seeded Gaussian beam through the planar
    undulator, uneven trajectories (8001-20001 points from random start z),
    16 requests of n x n points (n in 1..41, +-2 mm at 30 m) x 128-1024 energies
    (100-3000 eV), summed incoherently in one persistent launch per step (HIP
    graph: RK4 + spline + level samples + sums).  The FP64 tiled batch is the
    error reference; the plain FP64 path is timed on a particle subset."""
    import torch
    from oscars_amd import beam as B
    from oscars_amd import multiparticle as M
    from oscars_amd.sr import sr
    rng = np.random.Generator(np.random.PCG64(args.seed))
    beam = M.config_beam()
    parts = B.sample_particles(beam, args.particles, rng=rng)
    windows = M.uneven_windows(parts, rng)
    reqs = M.random_requests(16, rng)
    req_of, weights = M.assign_round_robin(args.particles, len(reqs))
    s = sr()
    s.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
    dev = s._device()
    print("# device: %s" % dev.info(), file=sys.stderr, flush=True)
    res = {}
    for mode in ([args.mode] if args.no_plain else [args.mode, "tiled"]):
        if mode in res:
            continue
        bf = M.BatchedFlux(dev, s._fields, parts, windows, reqs, req_of, weights, mode=mode)
        best, tt = bf.tune(candidates=(2, 4, 8, 16), sample_items=20000)
        print("# %s tuner: best %d waves/CU %s" % (mode, best, tt), file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            _, st = bf.run(recompute_trajectories=True)
            print("# %s warmup step %.1f ms" % (mode, st["kernel_ms"]), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        evals = 0.0
        kms = 0.0
        for _ in range(args.steps):
            outs, st = bf.run(recompute_trajectories=True)
            evals += st["evals"]
            kms += st["kernel_ms"]
            print("# %s step %.1f ms" % (mode, st["kernel_ms"]), file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[mode] = dict(value=evals / dt, ms=1e3 * dt / args.steps, evals=evals / args.steps, outs=outs,
                         unconverged=st["n_unconverged"], graph_ms=kms / args.steps, wpc=best)
        print("# %s: %.1f ms/step %.4g evals/s unconverged %d" % (mode, res[mode]["ms"], res[mode]["value"],
                                                                  st["n_unconverged"]), file=sys.stderr, flush=True)
    head = res[args.mode]
    extra = {}
    if "tiled" in res and args.mode != "tiled":
        extra["rel_err_vs_fp64_per_request"] = [rel_errors(a, b) for a, b in zip(head["outs"], res["tiled"]["outs"])]
    if not args.no_plain:
        # plain FP64 GPU path, particle by particle, on a subset of particles
        nsub = min(args.plain_particles, args.particles)
        dev.trajectories(s._fields, M.om_particles(parts[:nsub], windows[:nsub]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pe = 0.0
        for i in range(nsub):
            obs, e = reqs[req_of[i]]
            v, lev, cv, _, st = dev.flux(i, obs, e, [0.0] * 6, dev.conv(mode="plain"))
            pe += st["evals"]
        torch.cuda.synchronize()
        pdt = time.perf_counter() - t0
        extra["plain_fp64"] = {"value": pe / pdt, "particles": nsub}
        extra["ratio_vs_plain"] = head["value"] / (pe / pdt)
    line = {
        "metric": "LW integrand evals/s (obs x omega x traj), forward, batched multi-particle",
        "value": head["value"], "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64" if args.mode != "mixed" else "fp64 phase/amplitude, fp32 sincos", "data": "synthetic",
        "config": {"workload": "%d electrons (seed %d, Gaussian beam 0.9 nm/8 pm, beta 1.5/0.8 m, sigE/E 9e-4), "
                               "uneven 8001-20001-point trajectories, 16 requests n x n (n 1..41, +-2 mm @30 m) x "
                               "128-1024 energies 100-3000 eV, incoherent sum, HIP graph" % (args.particles, args.seed),
                   "mode": args.mode, "parallelism": "single GPU", "waves_per_cu": head["wpc"]},
        "evals_per_step": head["evals"], "unconverged": head["unconverged"],
    }
    line.update(extra)
    print(json.dumps(line, default=float))


def adjoint_bench(dev, steps=2):
    """Config: gradient of sum (F - F_ideal)^2 over a 51 x 51 grid (+-5 mm at 30 m) at
    E1 - 1.5, E1, 3 E1 - 4.5, 3 E1 (E1 = 152.11 eV) with respect to the 4001 samples of
    a By(z) map over +-1 m (planar undulator 1.0 T / 49 mm / 31 periods with 0.5 %
    per-pole amplitude errors and 20 um pole jitter, seed 5) and the beam angles,
    through the LW sum and the 20001-point RK4 trajectory.  The level is frozen at
    the highest level the error-free map converges at (precision 0.01)."""
    from oscars_amd import _lib
    from oscars_amd import adjoint as A
    from oscars_amd import geometry as G
    n, z0, dz = 4001, -1.0, 5e-4
    err = A.undulator_pole_map(n, z0, dz, sigma_amp=0.005, sigma_z=20e-6, seed=5)
    ideal = A.undulator_pole_map(n, z0, dz)
    obs = rect(51, 0.01)
    e1 = 152.11
    energies = np.array([e1 - 1.5, e1, 3 * e1 - 4.5, 3 * e1])
    f, keep = A.grid_field([0, 0, z0], [0, 0, dz], [1, 1, n], ideal)
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
    from oscars_amd import multiparticle as M
    dev.trajectories([f], M.om_particles([p0], [(0.0, 2.0, 20001)]))
    _, lev, _, _, _ = dev.flux(0, obs, energies, [0.0] * 6, dev.conv(mode="tiled"))
    L = int(lev.max())
    prob = A.FieldMapProblem(dev, [0, 0, z0], [0, 0, dz], [1, 1, n], 3.0, [0, 0, -1], 0.0, 2.0, 20001, obs, energies, L)
    f_ideal = prob.forward(ideal)
    tf = tb = 0.0
    for k in range(steps + 1):
        t0 = time.perf_counter()
        F = prob.forward(err)
        t1 = time.perf_counter()
        gf, ga, gx = prob.backward(2.0 * (F - f_ideal))
        t2 = time.perf_counter()
        if k > 0:  # first pass is warm-up
            tf += t1 - t0
            tb += t2 - t1
    tf /= steps
    tb /= steps
    evals = float(len(obs) * len(energies) * ((2 << L) - 1))
    return {"workload": "51x51 at 30 m x 4 energies, 4001-sample By map, 20001-point RK4, level %d" % L,
            "evals_per_pass": evals, "forward_ms": 1e3 * tf, "backward_ms": 1e3 * tb,
            "backward_kernel_ms": prob.last_stats["backward"]["kernel_ms"],
            "forward_evals_per_s": evals / tf, "adjoint_evals_per_s": evals / tb,
            "fwd_plus_adjoint_evals_per_s": 2 * evals / (tf + tb),
            "grad_norm_field": float(np.linalg.norm(gf)), "grad_angles": [float(v) for v in ga],
            "fd_check": "tests/test_gpu_adjoint.py (long-double central differences, 1e-4)"}


def batched_share(args, particles):
    """The multi-particle workload at `particles` electrons (the per-GPU share of the
    2048-electron job on an 8-GPU node at 256), mixed vs FP64 tiled."""
    import torch
    from oscars_amd import beam as B
    from oscars_amd import multiparticle as M
    from oscars_amd.sr import sr
    rng = np.random.Generator(np.random.PCG64(args.seed))
    parts = B.sample_particles(M.config_beam(), 2048, rng=rng)
    windows = M.uneven_windows(parts, rng)
    reqs = M.random_requests(16, rng)
    parts, windows = parts[:particles], windows[:particles]
    req_of, weights = M.assign_round_robin(particles, len(reqs))
    s = sr()
    s.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
    dev = s._device()
    out = {}
    res = {}
    for mode in ("mixed", "tiled"):
        bf = M.BatchedFlux(dev, s._fields, parts, windows, reqs, req_of, weights, mode=mode)
        best, tt = bf.tune(candidates=(2, 4, 8, 16), sample_items=20000)
        bf.run(recompute_trajectories=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs, st = bf.run(recompute_trajectories=True)
        dt = time.perf_counter() - t0
        res[mode] = outs
        out[mode] = {"evals_per_s": st["evals"] / dt, "ms_per_step": 1e3 * dt, "evals": st["evals"],
                     "unconverged": st["n_unconverged"], "rechecked_fp64": st["n_rechecked"],
                     "waves_per_cu": best}
        print("# batched %d %s: %.1f ms, %.4g evals/s" % (particles, mode, 1e3 * dt, st["evals"] / dt),
              file=sys.stderr, flush=True)
    out["rel_err_mixed_vs_fp64"] = max(rel_errors(a, b)["max_floored_1e3"] for a, b in zip(res["mixed"], res["tiled"]))
    out["particles"] = particles
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="mixed", choices=["mixed", "tiled", "plain"])
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--no-pipeline", action="store_true",
                    help="run the steps one after the other on one context (no trajectory/sums overlap)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--only-sweep", action="store_true",
                    help="time only the 51x51 x 512-energy sweep (for counter runs of the dominant kernel)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device-to-device copy and exit (reference for FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--workload", default="maps", choices=["maps", "batched"],
                    help="maps: single-electron flux maps + energy sweep (default); batched: multi-particle sums")
    ap.add_argument("--particles", type=int, default=2048)
    ap.add_argument("--plain-particles", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--batch-share", type=int, default=256,
                    help="particles of the multi-particle share reported beside the headline (0: skip)")
    ap.add_argument("--no-adjoint", action="store_true")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batched":
        return batched_bench(args)
    from oscars_amd.sr import sr
    s = sr()
    scenario(s)
    work = Workload(s)
    work.only_sweep = args.only_sweep
    works = [work]
    if not args.no_pipeline:
        s2 = sr()
        scenario(s2)
        works.append(Workload(s2))
        works[1].only_sweep = args.only_sweep
    print("# device: %s" % s._device().info(), file=sys.stderr, flush=True)

    dt, evals, kms, res = timed(works, args.mode, args.steps, args.warmup)
    ms_per_step = 1e3 * dt / args.steps
    value = evals / dt
    print("# %s: %.3f ms/step, %.4g evals/step, %.4g evals/s, kernel ms %s" %
          (args.mode, ms_per_step, evals / args.steps, value, kms), file=sys.stderr, flush=True)

    extra = {}
    sweep_ms = kms["sweep"]
    sweep_evals = res["sweep_evals"]
    n_obs_sweep = len(work.obs_sweep)
    # samples visited per observation point in the sweep: the tile kernel computes the amplitude once per
    # (obs, sample, 64-energy tile); count it per (obs, energy-tile, sample) actually visited.
    lev = res["sweep_levels"].reshape(n_obs_sweep, len(work.e_sweep))
    if args.mode == "plain":
        flops = FLOPS_POINT * sweep_evals
    else:
        tiles = [lev[:, i:i + 64].max(axis=1) for i in range(0, lev.shape[1], 64)]
        samples = float(sum(np.sum(2.0 ** (t + 1) - 1) for t in tiles))
        flops = FLOPS_TILE_EVAL * sweep_evals + FLOPS_TILE_SAMPLE * samples
    achieved = flops / (sweep_ms * 1e-3) / 1e12
    # HBM bytes per launch of the sweep kernel, measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    # on this workload and scaled by the counters' reading of a plain device copy (profiles/)
    traffic = None
    tj = {}
    tfile = os.path.join(ROOT, "profiles", "lw_tile_sweep_traffic.json")
    if os.path.exists(tfile) and args.mode == "mixed":
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": "lw_tile_kernel<%s>" % ("true" if args.mode == "mixed" else "false")
                if args.mode != "plain" else "lw_point_kernel<false>",
                "kernel_ms": round(sweep_ms, 3), "evals_per_launch": sweep_evals,
                "sincos_per_s": sweep_evals / (sweep_ms * 1e-3),
                "scope": "all launches of the sweep's tiled kernel (dense levels 0..8, then one per level)",
                "traffic_source": tj.get("source"),
                "valu_issue_frac": tj.get("valu_issue_frac"),
                "valu_wave_instr_per_eval": tj.get("valu_wave_instr_per_eval")}

    if len(works) == 2:
        # the same steps one after the other on one context (trajectory, then sums)
        sdt, sevals, _, _ = timed([work], args.mode, 2, 0)
        extra["sequential"] = {"ms_per_step": 1e3 * sdt / 2, "evals_per_s": sevals / sdt,
                               "note": "steps not pipelined: each step's RK4 runs before its sums"}
        print("# sequential: %.3f ms/step" % (1e3 * sdt / 2), file=sys.stderr, flush=True)

    if not args.no_plain and args.mode != "plain":
        pdt, pevals, pkms, pres = timed(works, "plain", args.plain_steps, 1 if len(works) == 2 else 0)
        pvalue = pevals / pdt
        extra["plain_fp64"] = {"value": pvalue, "ms_per_step": 1e3 * pdt / args.plain_steps, "kernel_ms": pkms}
        extra["ratio_vs_plain"] = value / pvalue
        errs = {}
        for k in res["values"]:
            errs[k] = rel_errors(res["values"][k], pres["values"][k])
            errs[k]["level_mismatches"] = int(np.sum(res["levels"][k] != pres["levels"][k]))
            errs[k]["points"] = int(res["levels"][k].size)
            errs[k]["rechecked_fp64"] = int(res["rechecked"][k])
        extra["rel_err_vs_fp64"] = errs
        print("# plain: %.4g evals/s ratio %.2f errs %s" % (pvalue, value / pvalue, errs), file=sys.stderr,
              flush=True)

    if not args.only_sweep:
        hms = host_api_step_ms(work, args.mode)
        extra["host_buffer_api"] = {"ms_per_step": hms, "evals_per_s": evals / args.steps / (hms * 1e-3),
                                    "note": "same step through om_flux (host grids and results, PCIe included)"}
        print("# host-buffer API: %.3f ms/step" % hms, file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        try:
            cb = cpu_baseline(work, res["sweep_levels"])
            map_evals = None
            # evals of the 152.11 eV map come from the GPU levels of the same points (identical levels)
            st_map = work.s._device().flux(0, work.obs_map, [152.11], [0.0] * 6,
                                           work.s._device().conv(mode="plain"))[4]
            map_evals = st_map["evals"]
            cpu_value = (map_evals + cb["sweep_evals"]) / (cb["t_map"] + cb["t_sweep"])
            cpu = {"value": cpu_value, "unit": "evals/s", "cores": cb["cores"], "kind": cb["kind"],
                   "sample": "201x201 map at 152.11 eV + 51x51 sweep at %d of the 512 energies, %.1f s" %
                             (cb["n_e"], cb["t_map"] + cb["t_sweep"])}
            print("# cpu baseline: %s" % cpu, file=sys.stderr, flush=True)
        except Exception as ex:  # the baseline is reported, never required
            cpu = {"value": None, "error": str(ex)}

    line = {
        "metric": "LW integrand evals/s (obs\u00d7\u03c9\u00d7traj) fwd+adjoint vs plain GPU at 1e-5 rel err",
        "value_is": "forward evals/s of the mixed path on the flux-map workload; adjoint and batched legs in "
                    "'adjoint' and 'batched', the plain FP64 GPU path in 'plain_fp64', errors in 'rel_err_vs_fp64'",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64" if args.mode != "mixed" else "fp64 phase/amplitude, fp32 sincos",
        "data": "synthetic",
        "config": {"workload": "planar undulator 1.0 T/49 mm/31 periods, 3 GeV 0.5 A, 20001-point RK4; "
                               "flux 201x201 (+-10 mm @30 m) at 152.11 and 260 eV + 51x51 x 512 energies "
                               "100-3000 eV", "mode": args.mode, "parallelism": "single GPU",
                   "steps_pipelined": len(works) == 2},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.only_sweep:
        if not args.no_adjoint:
            try:
                line["adjoint"] = adjoint_bench(s._device())
                print("# adjoint: %s" % line["adjoint"], file=sys.stderr, flush=True)
            except Exception as ex:
                line["adjoint"] = {"error": str(ex)}
        if args.batch_share > 0:
            try:
                line["batched"] = batched_share(args, args.batch_share)
            except Exception as ex:
                line["batched"] = {"error": str(ex)}
    line.update(extra)
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
