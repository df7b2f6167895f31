"""Benchmark: Lienard-Wiechert integrand evaluations per second on one MI355X.

Headline workload (one step): 2048 electrons sampled from a seeded Gaussian
beam (3 GeV, emittances 0.9 nm / 8 pm, beta 1.5 m / 0.8 m, alpha 0,
sigma_E/E 9e-4) through a planar undulator (By 1.0 T, period 49 mm, 31
periods); uneven trajectories of 8001-20001 RK4 points from random start z;
16 requests of n x n observation points (n 1..41, +-2 mm at 30 m) x 128-1024
energies (100-3000 eV), the particles summed incoherently per request.  A step
is one replay of the captured HIP graph: RK4 + spline of every particle, the
level sample rows, and all sums in one persistent launch, with reduced-accuracy
sin/cos (FP64 polynomials good to 4.8e-9) of the exactly reduced FP64 phase
(the mixed mode) and the full-accuracy recheck of borderline stop decisions.  One "eval" is one (observation point, energy,
trajectory sample) integrand evaluation; a point converged at level L costs
2^(L+1)-1 evals.

Reported beside the headline (one JSON line, rank 0, N=1):
  fp64_batched   the same step in FP64 (tiled kernel): per-request errors of the
                 mixed result against it, relative with a floor at 1e-6 of each
                 request's largest value and unfloored
  plain_fp64     the reference-faithful FP64 GPU path (one lane per output point,
                 FP64 sin/cos, particle after particle) on a seeded subset of the
                 particles; ratio_vs_plain = value / its evals/s
  adjoint        the field-map adjoint workload: flux on 51 x 51 points x 4
                 energies from a 4001-sample By(z) map, forward and reverse
  maps           the single-electron flux maps (201 x 201 at 152.11 and 260 eV,
                 51 x 51 x 512 energies), mixed and plain FP64
  power_density  power density on 401 x 401 points from an EPU trajectory,
                 mixed and FP64
  roofline       the persistent kernel's algorithmic flops over its own time
                 (HIP events around it on the library stream, eager launch)
  cpu_baseline   the reference's C++ core on the host cores, on a bounded sample
                 of the headline workload (median of 3)
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP64 flops (divisions and square roots count as one flop; an FMA as two):
#   per-point kernel: amplitude + phase 76, phase-step bookkeeping 3 per eval (sin/cos not counted)
#   batched kernel, per eval: phase 1, phase step 2, complex MAC 12, and sin/cos of the phase 27
#   (the reduced-accuracy evaluation: exact reduction modulo pi 7, the two polynomials 20);
#   plus the amplitude, 66 per (unit, level, sample) computed
FLOPS_POINT = 79.0
FLOPS_TILE_EVAL = 15.0
FLOPS_SINCOS = 27.0
FLOPS_TILE_AMP = 66.0
PEAK_FP64_TFLOPS = 78.6
METRIC = "LW integrand evals/s (obs×ω×traj) fwd+adjoint vs plain GPU at 1e-5 rel err"
CSRC = os.path.join(ROOT, "oscars_amd", "csrc")


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


# the sources the persistent batch kernel and its launch are built from
BATCH_SOURCES = ("batch.hip", "lw_common.h", "om_device.h", "om_types.h", "launch_types.h", "sincos_coeffs.h",
                 "rk4_common.h", "api.cpp")


def code_tokens(text):
    """C/C++ source without comments and whitespace (string literals kept)."""
    out, i, n = [], 0, len(text)
    while i < n:
        if text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        elif text[i] == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        else:
            if not text[i].isspace():
                out.append(text[i])
            i += 1
    return "".join(out)


def source_hash(read=None):
    """Hash of the batch kernel's code, comments and layout aside (ties a
    committed counter profile to the code it measured)."""
    read = read or (lambda name: open(os.path.join(CSRC, name)).read())
    h = hashlib.sha256()
    for name in BATCH_SOURCES:
        h.update(name.encode() + code_tokens(read(name)).encode())
    return h.hexdigest()[:16]


def rel_errors(a, b, floor_frac=1e-6):
    a = np.asarray(a, dtype=float).reshape(-1)
    b = np.asarray(b, dtype=float).reshape(-1)
    den = np.abs(b)
    pointwise = np.abs(a - b) / np.where(den > 0, den, 1.0)
    floored = np.abs(a - b) / np.maximum(den, floor_frac * np.max(den))
    return dict(max_floored_1e6=float(np.max(floored)), max_pointwise=float(np.max(pointwise)),
                l2=float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))


# ---------------------------------------------------------------- headline: batched multi-particle
class Config4:
    """The particles, windows and requests of the multi-particle workload (seeded)."""

    def __init__(self, n_particles, seed):
        from oscars_amd import beam as B
        from oscars_amd import multiparticle as M
        rng = np.random.Generator(np.random.PCG64(seed))
        self.beam = M.config_beam()
        self.parts = B.sample_particles(self.beam, n_particles, rng=rng)
        self.windows = M.uneven_windows(self.parts, rng)
        self.reqs = M.random_requests(16, rng)
        self.req_of, self.weights = M.assign_round_robin(n_particles, len(self.reqs))
        self.n = n_particles
        self.seed = seed

    def describe(self):
        return ("%d electrons (seed %d, Gaussian beam 3 GeV, emittance 0.9 nm / 8 pm, beta 1.5 / 0.8 m, "
                "sigma_E/E 9e-4), uneven 8001-20001-point RK4 trajectories, 16 requests n x n (n 1..41, +-2 mm "
                "@30 m) x 128-1024 energies 100-3000 eV, incoherent sum" % (self.n, self.seed))


def undulator_sr():
    from oscars_amd.sr import sr
    s = sr()
    s.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
    return s


def batched_headline(args, cfg):
    import torch
    from oscars_amd import multiparticle as M
    s = undulator_sr()
    dev = s._device()
    log("device: %s" % dev.info())
    t0 = time.perf_counter()
    bf = M.BatchedFlux(dev, s._fields, cfg.parts, cfg.windows, cfg.reqs, cfg.req_of, cfg.weights, mode="mixed")
    # the persistent grid is tuned on the whole workload: a prefix of the work
    # units is not representative of the mix of request sizes
    best, tt = bf.tune(candidates=(8, 16), sample_items=1 << 40)
    # then the level from which the points still running are handed over to
    # one workgroup each (0: none), on the whole launch
    deep, td = bf.tune_deep((0, 13, 15))
    log("setup + tuner %.1f s: %d waves/CU %s, hand-over level %d %s" % (time.perf_counter() - t0, best, tt, deep, td))
    for k in range(args.warmup):
        _, st = bf.run(recompute_trajectories=True)
        log("warmup step %d: %.1f ms" % (k, st["kernel_ms"]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evals = 0.0
    outs, st = None, None
    for k in range(args.steps):
        outs, st = bf.run(recompute_trajectories=True)
        evals += st["evals"]
        log("step %d: %.1f ms, %.4g evals, %d rechecked, %d unconverged" %
            (k, st["kernel_ms"], st["evals"], st["n_rechecked"], st["n_unconverged"]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = dict(value=evals / dt, ms=1e3 * dt / args.steps, evals_per_step=evals / args.steps, outs=outs,
               unconverged=st["n_unconverged"], rechecked=st["n_rechecked"], waves_per_cu=best,
               tuning={"waves_per_cu_ms": tt, "hand_over_level_ms": td, "hand_over_level": deep})
    # one more step launched stage by stage with HIP events around each kernel
    ms, cnt, pst = dev.batch_profile(True)
    res["profile"] = dict(stage_ms=ms, counts=cnt)
    log("profile: %s %s" % (ms, cnt))
    return s, dev, bf, res


def fp64_batched(dev, s, cfg, head):
    """The same step with the FP64 tiled batch kernel: the error reference."""
    import torch
    from oscars_amd import multiparticle as M
    bf = M.BatchedFlux(dev, s._fields, cfg.parts, cfg.windows, cfg.reqs, cfg.req_of, cfg.weights, mode="tiled")
    bf.run(recompute_trajectories=False)  # graph capture and first replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs, st = bf.run(recompute_trajectories=True)
    dt = time.perf_counter() - t0
    errs = [rel_errors(a, b) for a, b in zip(head["outs"], outs)]
    worst = {k: max(e[k] for e in errs) for k in errs[0]}
    log("fp64 batched: %.1f ms, %.4g evals/s; mixed vs fp64 worst %s" % (1e3 * dt, st["evals"] / dt, worst))
    return {"ms_per_step": 1e3 * dt, "evals_per_s": st["evals"] / dt, "evals": st["evals"],
            "evals_equal": st["evals"] == head["evals_per_step"],
            "rel_err_mixed_vs_fp64": worst, "rel_err_per_request": errs}


def plain_subset(dev, s, cfg, n_sub):
    """The plain FP64 GPU path (one lane per output point, FP64 sin/cos) on the
    first n_sub particles, one after the other, each on its own request, with
    their trajectories integrated on the device first (timed)."""
    import torch
    from oscars_amd import multiparticle as M
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.trajectories(s._fields, M.om_particles(cfg.parts[:n_sub], cfg.windows[:n_sub]))
    evals = 0.0
    conv = dev.conv(mode="plain")
    for i in range(n_sub):
        obs, e = cfg.reqs[cfg.req_of[i]]
        _, _, _, _, st = dev.flux(i, obs, e, [0.0] * 6, conv)
        evals += st["evals"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log("plain fp64 on %d particles: %.2f s, %.4g evals/s" % (n_sub, dt, evals / dt))
    return {"evals_per_s": evals / dt, "particles": n_sub, "evals": evals, "seconds": dt,
            "note": "particles 0..%d of the same sample (each on its own request), trajectories included" %
                    (n_sub - 1)}


def roofline(head, counters_file):
    """Algorithmic flops of the persistent kernel over its own time (HIP events
    around its eager launch on the library stream)."""
    ms = head["profile"]["stage_ms"]["main"]
    cnt = head["profile"]["counts"]
    flops = (FLOPS_TILE_EVAL + FLOPS_SINCOS) * cnt["main_evals"] + FLOPS_TILE_AMP * cnt["main_amplitudes"]
    flops_nosc = FLOPS_TILE_EVAL * cnt["main_evals"] + FLOPS_TILE_AMP * cnt["main_amplitudes"]
    achieved = flops / (ms * 1e-3) / 1e12
    traffic, note = None, None
    extra = {}
    if os.path.exists(counters_file):
        with open(counters_file) as f:
            tj = json.load(f)
        if tj.get("source_hash") == source_hash():
            traffic = tj.get("hbm_bytes_per_launch")
            extra = {k: tj.get(k) for k in ("valu_issue_frac", "valu_wave_instr_per_eval", "hbm_gbs", "clock_ghz",
                                            "fp64_tflops_counted")}
            note = tj.get("source")
        else:
            note = "%s measured other kernel sources (hash %s, now %s): traffic not reported" % (
                os.path.relpath(counters_file, ROOT), tj.get("source_hash"), source_hash())
    out = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
           "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
           "frac_without_sincos": round(flops_nosc / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
           "lane_utilization": round(cnt["main_evals"] / max(cnt["main_slots"], 1), 4),
           "kernel": "lw_batch_kernel<true>", "kernel_ms": round(ms, 3),
           "evals_per_launch": cnt["main_evals"], "amplitudes_per_launch": cnt["main_amplitudes"],
           "flops_per_launch": flops, "sincos_per_s": cnt["main_evals"] / (ms * 1e-3),
           "stage_ms": head["profile"]["stage_ms"], "counters_source": note}
    out.update(extra)
    return out


# ---------------------------------------------------------------- CPU baseline
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(dev, s, cfg, n_parts=6, n_e=6, repeats=3):
    """The reference's compiled C++ core (oracle/_ref/libref.so, built from the
    reference sources; the C restatement, kind 'port', where it is absent) on
    the host cores, on a bounded sample of the headline workload: particles
    0..n_parts-1 with n_e evenly spaced energies of their requests, each a
    calculate_flux_rectangle call of the reference on the request's grid.
    Evals are counted with the GPU plain path's per-point levels for the same
    points (the plain path's levels equal the reference's)."""
    from oracle import pyoracle as P
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    kind = "reference" if P.ref_available() else "port"
    from oscars_amd import multiparticle as M
    jobs, evals = [], 0.0
    dev.trajectories(s._fields, M.om_particles(cfg.parts[:n_parts], cfg.windows[:n_parts]))
    conv = dev.conv(mode="plain")
    for i in range(n_parts):
        obs, e = cfg.reqs[cfg.req_of[i]]
        n = int(round(np.sqrt(len(obs))))
        es = e[np.linspace(0, len(e) - 1, n_e).round().astype(int)]
        _, _, _, _, st = dev.flux(i, obs, es, [0.0] * 6, conv)
        evals += st["evals"]
        jobs.append((cfg.parts[i], cfg.windows[i], n, es))
    times = []
    for _ in range(repeats):
        t = 0.0
        for p, (a, b, npts), n, es in jobs:
            sim = P.RefSim(nthreads=ncores) if kind == "reference" else P.Oracle(nthreads=ncores)
            sim.set_particle_beam(energy_GeV=3.0, x0=p.x0, d0=[0, 0, 1], current=0.5)
            sim.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
            sim.set_particle(p.x0, p.b0, p.t0)
            sim.set_ctstartstop(a, b)
            sim.set_npoints_trajectory(npts)
            t0 = time.perf_counter()
            for en in es:
                sim.calculate_flux_rectangle(float(en), [n, n], [0.004, 0.004], translation=[0, 0, 30])
            t += time.perf_counter() - t0
        times.append(t)
    tm = float(np.median(times))
    log("cpu baseline (%s, %d cores, %s): %s s -> %.4g evals/s" % (kind, ncores, cpu_model(), times, evals / tm))
    return {"value": evals / tm, "unit": "evals/s", "cores": ncores, "kind": kind,
            "sample": "particles 0..%d of the headline sample, %d energies of each one's request on its full grid "
                      "(%.3g evals, %.2f s median of %d runs, trajectory integration included)" %
                      (n_parts - 1, n_e, evals, tm, repeats),
            "cpu_model": cpu_model(), "runs_s": times}


# ---------------------------------------------------------------- adjoint leg
def rect(n, w=0.02):
    from oscars_amd import geometry as G
    obs, nv, cont = G.rectangle("XY", n, n, w, w, [0, 0, 0], [0, 0, 30], 0, 2)
    return np.array(obs)


def adjoint_leg(dev, steps=2):
    """Gradient of sum (F - F_ideal)^2 over a 51 x 51 grid (+-5 mm at 30 m) at
    E1 - 1.5, E1, 3 E1 - 4.5, 3 E1 (E1 = 152.11 eV) with respect to the 4001
    samples of a By(z) map over +-1 m (planar undulator 1.0 T / 49 mm / 31
    periods with 0.5 % per-pole amplitude errors and 20 um pole jitter, seed 5)
    and the beam angles, through the LW sum and the 20001-point RK4 trajectory.
    The level is frozen at the highest level the error-free map converges at
    (precision 0.01)."""
    from oscars_amd import adjoint as A
    from oscars_amd import geometry as G
    from oscars_amd import multiparticle as M
    n, z0, dz = 4001, -1.0, 5e-4
    err = A.undulator_pole_map(n, z0, dz, sigma_amp=0.005, sigma_z=20e-6, seed=5)
    ideal = A.undulator_pole_map(n, z0, dz)
    obs = rect(51, 0.01)
    e1 = 152.11
    energies = np.array([e1 - 1.5, e1, 3 * e1 - 4.5, 3 * e1])
    f, keep = A.grid_field([0, 0, z0], [0, 0, dz], [1, 1, n], ideal)
    p0 = G.ideal_particle("electron", 3.0, [0, 0, -1], [0, 0, 1], 0.0, 0.5)
    dev.trajectories([f], M.om_particles([p0], [(0.0, 2.0, 20001)]))
    _, lev, _, _, _ = dev.flux(0, obs, energies, [0.0] * 6, dev.conv(mode="tiled"))
    L = int(lev.max())
    import torch
    prob = A.FieldMapProblem(dev, [0, 0, z0], [0, 0, dz], [1, 1, n], 3.0, [0, 0, -1], 0.0, 2.0, 20001, obs, energies, L)
    # samples, flux and gradients resident on the device (om_adjoint_*_dev)
    err_t = torch.tensor(err, dtype=torch.float64, device="cuda")
    f_ideal = prob.forward_dev(torch.tensor(ideal, dtype=torch.float64, device="cuda")).clone()
    tf = tb = 0.0
    for k in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F = prob.forward_dev(err_t)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gf, ga, gx = prob.backward_dev(2.0 * (F - f_ideal))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if k > 0:  # the first pass is warm-up
            tf += t1 - t0
            tb += t2 - t1
    tf /= steps
    tb /= steps
    gf = gf.cpu().numpy()
    ga = ga.cpu().numpy()
    # for scale: the plain FP64 forward of the same problem (RK4 of the error
    # map's trajectory + one lane per point summing levels 0..L), device tensors
    fe, keep_e = A.grid_field([0, 0, z0], [0, 0, dz], [1, 1, n], err)
    parts = M.om_particles([p0], [(0.0, 2.0, 20001)])
    obs_t = torch.tensor(obs, dtype=torch.float64, device="cuda")
    e_t = torch.tensor(energies, dtype=torch.float64, device="cuda")
    out_t = torch.zeros(len(obs) * len(energies), dtype=torch.float64, device="cuda")
    conv = dev.conv(precision=0.0, max_level=L, mode="plain")
    tp = 0.0
    for k in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.trajectories([fe], parts)
        dev.flux_dev(0, obs_t, e_t, [0.0] * 6, conv, out_t)
        torch.cuda.synchronize()
        if k > 0:
            tp += time.perf_counter() - t0
    tp /= steps
    plain_err = rel_errors(out_t.cpu().numpy(), F.reshape(-1).cpu().numpy())
    evals = float(len(obs) * len(energies) * ((2 << L) - 1))
    out = {"workload": "51x51 at 30 m x 4 energies, 4001-sample By map, 20001-point RK4, level %d; samples, flux "
                       "and gradients as device tensors" % L,
           "evals_per_pass": evals, "forward_ms": 1e3 * tf, "backward_ms": 1e3 * tb,
           "forward_kernel_ms": prob.last_stats["forward"]["kernel_ms"],
           "backward_kernel_ms": prob.last_stats["backward"]["kernel_ms"],
           "forward_evals_per_s": evals / tf, "adjoint_evals_per_s": evals / tb,
           "fwd_plus_adjoint_evals_per_s": 2 * evals / (tf + tb),
           "plain_forward_ms": 1e3 * tp, "plain_forward_evals_per_s": evals / tp,
           "fwd_plus_adjoint_vs_plain_forward": (2 * evals / (tf + tb)) / (evals / tp),
           "plain_forward_vs_adjoint_forward_rel_err": plain_err,
           "grad_norm_field": float(np.linalg.norm(gf)), "grad_angles": [float(v) for v in ga],
           "fd_check": "tests/test_gpu_adjoint.py (long-double central differences, 1e-4)"}
    log("adjoint: forward %.1f ms, backward %.1f ms; plain FP64 forward of the same problem %.1f ms" %
        (1e3 * tf, 1e3 * tb, 1e3 * tp))
    return out


# ---------------------------------------------------------------- single-electron maps leg
def map_scenario(s):
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1], d0=[0, 0, 1], current=0.5)
    s.add_bfield_undulator(bfield=[0, 1, 0], period=[0, 0, 0.049], nperiods=31)
    s.set_ctstartstop(0, 2)
    s.set_npoints_trajectory(20001)


def maps_leg(steps=4):
    """Flux 201 x 201 (+-10 mm at 30 m) at 152.11 and 260 eV and 51 x 51 x 512
    energies (100-3000 eV) on a 20001-point trajectory; device-resident inputs
    and outputs; every step integrates its trajectory first (no pipelining).
    The maps run on the FP64 per-point kernel, the sweep mixed; the plain path
    runs everything one lane per point in FP64."""
    import torch
    from oscars_amd.sr import sr
    s = sr()
    map_scenario(s)
    dev = s._device()
    cuda = dict(device="cuda")
    obs_map, obs_sweep = rect(201), rect(51)
    e_sweep = np.array([100 + (3000 - 100) / 511 * i for i in range(512)])
    t_map = torch.tensor(obs_map, dtype=torch.float64, **cuda)
    t_sweep = torch.tensor(obs_sweep, dtype=torch.float64, **cuda)
    t_es = torch.tensor(e_sweep, dtype=torch.float64, **cuda)
    t_em = {e: torch.tensor([e], dtype=torch.float64, **cuda) for e in (152.11, 260.0)}
    outs = {}
    for k, n in (("map_152.11", len(obs_map)), ("map_260", len(obs_map)), ("sweep", len(obs_sweep) * 512)):
        outs[k] = (torch.zeros(n, dtype=torch.float64, **cuda), torch.zeros(n, dtype=torch.int32, **cuda),
                   torch.zeros(n, dtype=torch.uint8, **cuda))

    calls = {}

    def step(mode):
        dev.trajectories(s._fields, [s._om_particle(s._particle)], check_status=False)
        ev, kms = 0.0, {}
        for e in (152.11, 260.0):
            k = "map_%g" % e
            st = dev.flux_dev(0, t_map, t_em[e], [0.0] * 6, dev.conv(mode="plain"), *outs[k])
            ev += st["evals"]
            kms[k] = st["kernel_ms"]
            calls[k] = st["evals"]
        st = dev.flux_dev(0, t_sweep, t_es, [0.0] * 6, dev.conv(mode=mode), *outs["sweep"])
        ev += st["evals"]
        kms["sweep"] = st["kernel_ms"]
        calls["sweep"] = st["evals"]
        return ev, kms

    res = {}
    vals = {}
    for mode in ("mixed", "plain"):
        step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev = 0.0
        for _ in range(steps):
            e, kms = step(mode)
            ev += e
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[mode] = {"ms_per_step": 1e3 * dt / steps, "evals_per_s": ev / dt, "kernel_ms": kms,
                     "evals_per_call": dict(calls)}
        vals[mode] = {k: (v[0].cpu().numpy(), v[1].cpu().numpy()) for k, v in outs.items()}
        log("maps %s: %.2f ms/step, %.4g evals/s" % (mode, 1e3 * dt / steps, ev / dt))
    errs = {}
    for k in vals["mixed"]:
        errs[k] = rel_errors(vals["mixed"][k][0], vals["plain"][k][0])
        errs[k]["level_mismatches"] = int(np.sum(vals["mixed"][k][1] != vals["plain"][k][1]))
    res["ratio_mixed_vs_plain"] = res["mixed"]["evals_per_s"] / res["plain"]["evals_per_s"]
    res["rel_err_mixed_vs_plain"] = errs
    res["workload"] = ("one electron, 20001-point RK4; flux 201x201 (+-10 mm @30 m) at 152.11 and 260 eV + "
                       "51x51 x 512 energies 100-3000 eV")
    return res


# ---------------------------------------------------------------- power-density leg
def epu_scenario(s):
    """Elliptically polarising undulator: 0.6 T x 0.6 T, period 57 mm, 40 periods,
    3 GeV from z = -1.5 m, 40001-point trajectory over ct 0..3 m."""
    s.set_particle_beam(energy_GeV=3, x0=[0, 0, -1.5], d0=[0, 0, 1], current=0.5)
    s.add_bfield_epu(bfieldA=[0, 0.6, 0], bfieldB=[0.6, 0, 0], period=[0, 0, 0.057], nperiods=40,
                     phase=np.pi / 2)
    s.set_ctstartstop(0, 3)
    s.set_npoints_trajectory(40001)


def power_density_leg(steps=4):
    """Power density on 401 x 401 points (+-40 mm at 20 m, normal +z) from the
    EPU trajectory, device-resident inputs and outputs, mixed and FP64; every
    step integrates its trajectory first."""
    import torch
    from oscars_amd import geometry as G
    from oscars_amd.sr import sr
    s = sr()
    epu_scenario(s)
    dev = s._device()
    obs, _, _ = G.rectangle("XY", 401, 401, 0.08, 0.08, [0, 0, 0], [0, 0, 20.0], 0, 2)
    t_obs = torch.tensor(np.array(obs), dtype=torch.float64, device="cuda")
    n = t_obs.shape[0]
    out = {m: (torch.zeros(n, dtype=torch.float64, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda"),
               torch.zeros(n, dtype=torch.uint8, device="cuda")) for m in ("mixed", "plain")}

    def step(mode):
        dev.trajectories(s._fields, [s._om_particle(s._particle)], check_status=False)
        return dev.power_density_dev(0, t_obs, [0, 0, 1], 0, dev.conv(mode=mode), *out[mode])

    res = {}
    for mode in ("mixed", "plain"):
        step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms, ev = 0.0, 0.0
        for _ in range(steps):
            st = step(mode)
            kms += st["kernel_ms"]
            ev += st["evals"]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[mode] = {"ms_per_step": 1e3 * dt / steps, "kernel_ms": kms / steps, "evals_per_step": ev / steps,
                     "kernel_evals_per_s": ev / (kms * 1e-3), "rechecked_fp64": st["n_rechecked"]}
        log("power density %s: %.2f ms/step, power-density kernels %.3f ms" % (mode, 1e3 * dt / steps, kms / steps))
    vm, lm = out["mixed"][0].cpu().numpy(), out["mixed"][1].cpu().numpy()
    vp, lp = out["plain"][0].cpu().numpy(), out["plain"][1].cpu().numpy()
    res["rel_err_mixed_vs_plain"] = rel_errors(vm, vp)
    res["level_mismatches"] = int(np.sum(lm != lp))
    res["kernel_ratio_mixed_vs_plain"] = res["plain"]["kernel_ms"] / res["mixed"]["kernel_ms"]
    res["stl"] = stl_leg()
    res["workload"] = ("EPU 0.6 T x 0.6 T, 40 x 57 mm, 3 GeV, 40001-point RK4; power density 401x401 "
                       "(+-40 mm @20 m)")
    return res


def stl_leg(subdiv=4):
    """Power density on the centres of a 5120-triangle sphere (radius 10 mm at
    z = 5 m, the layout of the reference's CAD example) with occlusion, from
    the planar undulator's trajectory: the bounding-volume hierarchy against
    testing every triangle for every ray (the reference's loop)."""
    from oscars_amd import geometry as G
    from oscars_amd.sr import sr
    s = sr()
    map_scenario(s)
    s._ensure_trajectory()
    dev = s._device()
    tri = G.stl_rows(G.icosphere(0.010, subdiv, (0.0, 0.0, 5.0)))
    out = {}
    for mode in ("tiled", "plain"):
        dev.power_density_stl(0, tri, dev.conv(mode=mode))  # warm-up
        v, lev, _, st = dev.power_density_stl(0, tri, dev.conv(mode=mode))
        out[mode] = (v, lev, st)
        log("stl %s: %d triangles, kernels %.2f ms" % (mode, len(tri), st["kernel_ms"]))
    return {"triangles": len(tri), "bvh_kernel_ms": out["tiled"][2]["kernel_ms"],
            "every_triangle_kernel_ms": out["plain"][2]["kernel_ms"],
            "speedup": out["plain"][2]["kernel_ms"] / out["tiled"][2]["kernel_ms"],
            "identical": bool(np.array_equal(out["tiled"][0], out["plain"][0]) and
                              np.array_equal(out["tiled"][1], out["plain"][1])),
            "workload": "planar undulator 1 T / 49 mm / 31 periods, 20001-point RK4; sphere of 10 mm at z = 5 m, "
                        "%d triangles" % len(tri)}


# ---------------------------------------------------------------- counter runs
def counters_run(args, cfg):
    """Setup, then one eager batch step (for rocprofv3 --pmc runs of the persistent kernel)."""
    from oscars_amd import multiparticle as M
    s = undulator_sr()
    dev = s._device()
    bf = M.BatchedFlux(dev, s._fields, cfg.parts, cfg.windows, cfg.reqs, cfg.req_of, cfg.weights, mode="mixed",
                       waves_per_cu=args.waves_per_cu)
    ms, cnt, _ = dev.batch_profile(True)
    print(json.dumps({"stage_ms": ms, "counts": cnt, "source_hash": source_hash()}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain-particles", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-maps", action="store_true")
    ap.add_argument("--no-pd", action="store_true")
    ap.add_argument("--power-density-only", action="store_true",
                    help="run the power-density leg alone and print its JSON (for rocprofv3 runs)")
    ap.add_argument("--maps-only", action="store_true",
                    help="run the single-electron maps leg alone and print its JSON (for rocprofv3 runs)")
    ap.add_argument("--counters", action="store_true",
                    help="one eager batch step after setup, nothing timed (for rocprofv3 --pmc runs)")
    ap.add_argument("--waves-per-cu", type=int, default=16, help="persistent grid of a --counters run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device-to-device copy and exit (reference for FETCH_SIZE/WRITE_SIZE)")
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
    if args.power_density_only:
        print(json.dumps(power_density_leg(), default=float))
        return
    if args.maps_only:
        print(json.dumps(maps_leg(), default=float))
        return
    cfg = Config4(args.particles, args.seed)
    if args.counters:
        return counters_run(args, cfg)
    s, dev, bf, head = batched_headline(args, cfg)
    log("headline: %.1f ms/step, %.4g evals/step, %.4g evals/s" % (head["ms"], head["evals_per_step"],
                                                                   head["value"]))
    line = {
        "metric": METRIC,
        "value_is": "forward evals/s of the mixed-precision batched multi-particle path (one persistent launch "
                    "per step in a HIP graph); the adjoint leg in 'adjoint', the plain FP64 GPU path in "
                    "'plain_fp64' and 'ratio_vs_plain', errors against the FP64 batch in 'fp64_batched'",
        "value": head["value"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 (sin/cos to 4.8e-9, borderline points rechecked at full accuracy)",
        "data": "synthetic",
        "config": {"workload": cfg.describe(), "mode": "mixed", "parallelism": "single GPU",
                   "waves_per_cu": head["waves_per_cu"], "tuning": head["tuning"],
                   "evals_per_step": head["evals_per_step"],
                   "unconverged": head["unconverged"], "rechecked_fp64": head["rechecked"]},
        "roofline": roofline(head, os.path.join(ROOT, "profiles", "lw_batch_counters.json")),
        "cpu_baseline": None,
    }
    if not args.no_fp64:
        try:
            line["fp64_batched"] = fp64_batched(dev, s, cfg, head)
        except Exception as ex:  # reported, never required
            line["fp64_batched"] = {"error": str(ex)}
    if not args.no_plain:
        try:
            pl = plain_subset(dev, s, cfg, min(args.plain_particles, args.particles))
            line["plain_fp64"] = pl
            line["ratio_vs_plain"] = head["value"] / pl["evals_per_s"]
        except Exception as ex:
            line["plain_fp64"] = {"error": str(ex)}
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(dev, s, cfg)
        except Exception as ex:
            line["cpu_baseline"] = {"value": None, "error": str(ex)}
    if not args.no_adjoint:
        try:
            line["adjoint"] = adjoint_leg(dev)
        except Exception as ex:
            line["adjoint"] = {"error": str(ex)}
    if not args.no_maps:
        try:
            line["maps"] = maps_leg()
        except Exception as ex:
            line["maps"] = {"error": str(ex)}
    if not args.no_pd:
        try:
            line["power_density"] = power_density_leg()
        except Exception as ex:
            line["power_density"] = {"error": str(ex)}
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
