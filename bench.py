#!/usr/bin/env python
"""CSR step benchmark on one MI355X.

Workload (default ``large``): the large single run -- 1e7 macroparticles,
1 nC, 2 GeV, two-Gaussian longitudinal profile (0.3 of the charge with
sigma 40 um at z = -75 um, 0.7 with sigma 80 um at +75 um), sigma_x = 20 um,
through a 4-bend chicane (0.3 m bends at +, -, -, + 0.06 rad, drifts 2/1/2 m)
in 1000 steps; deposition 256 x 256 with Savitzky-Golay 9/3, wake mesh
128 x 128, integration 200 x 200, n_formation_length 4, plain float64, the
whole history on the device.  Also ``chicane`` (1e6 particles, 0.05 rad,
400 steps, deposition 128 x 128 SG 7/2, mesh 64 x 64) and ``dipole`` (2e5
particles, one 0.5 m bend, 100 steps, deposition 64 x 64, mesh 32 x 32).

A step is one lattice step of the simulation: tracking plus one CSR step
(deposition -> density functions -> history -> retarded integral -> kicks).
Before timing, the run is filled to its steady state (FILL steps, 450 for
the large run: the retarded-time window is full, 430 slices on 500 x 500,
and slices are being dropped), then --warmup more untimed steps, then
exactly --steps timed ones.  ``value`` is the integrand throughput of the
timed steps: nominal integrand points (one point = one (x', s') source for
one mesh point; M x P per step) per second of the retarded integral.
``points_evaluated`` counts the points the kernels actually evaluate (the
others lie outside the history, where the reference's integrand is exactly
0, or gather a density and gradients that the history's span table shows to
be 0, which gives exactly 0 too: ``points_zero_skipped``); ``roofline``
prices the evaluated points at the cost model's 214 FP64 ops per point.

The JSON line also carries the roofline of the dominant work (the wake
integral: its four launches timed together with HIP events on the stream
they run on; traffic from the committed counter runs), the mixed-precision
form on the same state with its error and its own roofline, and CPU
baselines: the oracle's C/OpenMP restatement of the same integral on all
host cores over a bounded sample of mesh points, and its numpy restatement
on one core over a smaller sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from collections import OrderedDict

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CSR step ms + integrand Gpts/s, mixed vs FP64 @1e-4 rel err; serve p50/p99"
FP64_PEAK_TFLOPS = 78.6     # MI355X vector/matrix FP64 peak
FP32_PEAK_TFLOPS = 157.3    # MI355X vector FP32 peak (packed FP32)
FLOPS_PER_POINT = 214       # ops per integrand point of the reference formula
# the mixed form's share of those ops kept in float64: lattice lookups (22),
# |r - r'| and geometry (12), t_ret and z' (2), 3D index math (12), the
# trapezoid accumulation (5); the rest (161) in float32
MIXED_FP64_OPS = 53


def chicane_config(n_particle=1_000_000, nbins=128, window=7, mesh=64, Nxz=200):
    lat = OrderedDict(step_size=6.2 / 400)
    lat["B1"] = dict(type="dipole", L=0.3, angle=0.05, E1=0.0, E2=0.05, nsep=1)
    lat["D1"] = dict(type="drift", L=2.0, nsep=1)
    lat["B2"] = dict(type="dipole", L=0.3, angle=-0.05, E1=-0.05, E2=0.0, nsep=1)
    lat["D2"] = dict(type="drift", L=1.0, nsep=1)
    lat["B3"] = dict(type="dipole", L=0.3, angle=-0.05, E1=0.0, E2=-0.05, nsep=1)
    lat["D3"] = dict(type="drift", L=2.0, nsep=1)
    lat["B4"] = dict(type="dipole", L=0.3, angle=0.05, E1=0.05, E2=0.0, nsep=1)
    return {
        "input_beam": dict(style="gaussian", n_particle=n_particle, charge=1e-9, energy=1e9,
                           sigma_x=30e-6, sigma_z=100e-6, chirp=-10.0, sigma_delta=1e-4, seed=0),
        "input_lattice": dict(lattice_input_file="<chicane>", _config=lat),
        "particle_deposition": dict(xbins=nbins, zbins=nbins, xlim=5, zlim=5, filter_order=2,
                                    filter_window=window, velocity_threhold=1000),
        "CSR_integration": dict(n_formation_length=4, zbins=Nxz, xbins=Nxz),
        "CSR_computation": dict(xbins=mesh, zbins=mesh, xlim=5, zlim=5, apply_CSR=1,
                                transverse_on=1, write_wakes=False),
    }


def dipole_config(n_particle=200_000, nbins=64, window=5, mesh=32, Nxz=200):
    """Single dipole, plain method: 2e5 macroparticles, 1 nC, 5 GeV, Gaussian
    sigma_z = sigma_x = 50 um, no transverse momentum or energy spread, seed 0;
    rho = 10 m, L = 0.5 m (0.05 rad), 100 steps; deposition 64 x 64 over
    +/-5 sigma with Savitzky-Golay 5/2; wake mesh 32 x 32; integration
    200 x 200, n_formation_length 4 (the reference's defaults)."""
    lat = OrderedDict(step_size=0.5 / 100)
    lat["B1"] = dict(type="dipole", L=0.5, angle=0.05, E1=0.0, E2=0.0, nsep=1)
    return {
        "input_beam": dict(style="gaussian", n_particle=n_particle, charge=1e-9, energy=5e9,
                           sigma_x=50e-6, sigma_z=50e-6, chirp=0.0, sigma_delta=0.0, seed=0),
        "input_lattice": dict(lattice_input_file="<dipole>", _config=lat),
        "particle_deposition": dict(xbins=nbins, zbins=nbins, xlim=5, zlim=5, filter_order=2,
                                    filter_window=window, velocity_threhold=1000),
        "CSR_integration": dict(n_formation_length=4, zbins=Nxz, xbins=Nxz),
        "CSR_computation": dict(xbins=mesh, zbins=mesh, xlim=5, zlim=5, apply_CSR=1,
                                transverse_on=1, write_wakes=False),
    }


def large_config(n_particle=10_000_000, nbins=256, window=9, order=3, mesh=128, Nxz=200):
    """Large single run: 1e7 macroparticles, 1 nC, 2 GeV; two-Gaussian z
    profile (fraction 0.3 with sigma 40 um centred at z = -75 um, fraction 0.7
    with sigma 80 um at z = +75 um, z > 0 the bunch head), sigma_x = 20 um,
    no chirp or energy spread, seed 0; 4-bend chicane (0.3 m bends at
    +, -, -, + 0.06 rad, drifts 2/1/2 m), 1000 steps; deposition 256 x 256
    with Savitzky-Golay 9/3; wake mesh 128 x 128; integration 200 x 200."""
    cfg = chicane_config(n_particle=n_particle, nbins=nbins, window=window, mesh=mesh, Nxz=Nxz)
    lat = cfg["input_lattice"]["_config"]
    lat["step_size"] = 6.2 / 1000
    for name, sign in (("B1", 1), ("B2", -1), ("B3", -1), ("B4", 1)):
        e = lat[name]
        e["angle"] = sign * 0.06
        e["E1"] = 0.0 if e["E1"] == 0.0 else sign * 0.06
        e["E2"] = 0.0 if e["E2"] == 0.0 else sign * 0.06
    cfg["input_beam"] = dict(style="gaussian", n_particle=n_particle, charge=1e-9, energy=2e9,
                             sigma_x=20e-6, sigma_z=60e-6, chirp=0.0, sigma_delta=0.0, seed=0,
                             two_gaussian=dict(sigmas=(40e-6, 80e-6), separation=150e-6,
                                               ratio=(0.3, 0.7)))
    cfg["particle_deposition"]["filter_order"] = order
    return cfg


WORKLOADS = {
    # name: (config, fill steps to the steady state, description)
    "large": (large_config, 450,
              "large chicane run: 1e7 particles, 1 nC, 2 GeV, two-Gaussian z (0.3 x 40 um at "
              "-75 um, 0.7 x 80 um at +75 um), sigma_x 20 um, 4 bends x 0.3 m at 0.06 rad, "
              "1000 steps; deposition 256x256 SG 9/3; wake mesh 128x128; integration "
              "200x200; n_fl 4"),
    "chicane": (chicane_config, 200,
                "chicane bunch compressor: 1e6 particles, 1 nC, 1 GeV, 4 bends x 0.3 m at "
                "0.05 rad, drifts 2/1/2 m, 400 steps; deposition 128x128 SG 7/2; wake mesh "
                "64x64; integration 200x200; n_fl 4"),
    "dipole": (dipole_config, 90,
               "single dipole: 2e5 particles, 1 nC, 5 GeV, rho 10 m, L 0.5 m, 100 steps; "
               "deposition 64x64 SG 5/2; wake mesh 32x32; integration 200x200; n_fl 4"),
}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


class _HostHistory:
    """The device history's five fields (nt, Xi, Zi), in logical order, on
    the host, with the axes the oracle's lookups read."""

    def __init__(self, tr):
        self.data = {k: v.cpu().numpy() for k, v in tr.history_fields().items()}
        self._axes = dict(min_t=float(tr.min_x), min_x=float(tr.min_y), min_z=float(tr.min_z),
                          dt=float(tr.delta_x), dx=float(tr.delta_y), dz=float(tr.delta_z))

    def axes(self):
        return self._axes


def cpu_baseline(sim, gpu_out, budget_s=15.0, numpy_budget_s=8.0, max_points=None):
    """The oracle's C/OpenMP integral (all host cores) and its numpy
    restatement (one core) on the simulation's current history and wake
    parameters, each on the leading mesh points that fit its time budget;
    gpu_out is the FP64 device wake for the same inputs."""
    import numpy as np
    from oracle import cpu_ref as O
    from oracle import csr_ref_c as C
    tr = sim.DF_tracker
    hh = _HostHistory(tr)
    keep = C._Keep()
    H = C.hist_from_tracker(hh, keep)
    L = C.lat_from_tables(sim.lattice, keep)
    b = sim.beam
    beam = dict(sigma_x=b._sigma_x, sigma_z=b._sigma_z, slope=b._slope, mean_x=b._mean_x)
    P = sim.wake_params()
    obs = np.stack([gpu_out["zmesh"], gpu_out["xmesh"]], axis=1)
    ip = sim.integration_params
    per_point = sim.integrand_points() // sim._mesh_M()
    threads = C.threads()
    gpu_dE = gpu_out["dE_dct"]
    # probe, then size the sample to the time budget
    n0 = min(threads, obs.shape[0])
    t0 = time.perf_counter()
    C.wake(H, L, P.t, beam, P.formation_length, P.n_formation_length, ip.xbins, ip.zbins,
           P.csr_scaling, obs[:n0], threads)
    dt0 = time.perf_counter() - t0
    k = int(min(obs.shape[0], max(n0, n0 * budget_s / max(dt0, 1e-6))))
    if max_points:
        k = min(k, max_points)
    k = max(threads * (k // threads), n0)
    t0 = time.perf_counter()
    dE, _ = C.wake(H, L, P.t, beam, P.formation_length, P.n_formation_length, ip.xbins, ip.zbins,
                   P.csr_scaling, obs[:k], threads)
    dt = time.perf_counter() - t0
    err = rel_err(gpu_dE[:k], dE)[0]
    model = cpu_model()
    res = {"value": k * per_point / dt / 1e9, "unit": "Gpts/s", "cores": threads,
           "kind": "port", "cpu_model": model,
           "sample": f"oracle C/OpenMP integral (every nominal point evaluated, as the "
                     f"reference does) on {k} of {obs.shape[0]} mesh points of the last timed "
                     f"step ({per_point} integrand points each, {dt:.1f} s)",
           "max_rel_diff_vs_gpu_dE": err}
    # numpy restatement on one core (the reference's serial numpy/numba loop)
    lat = O.LatticeTables([dict(v) for v in sim.lattice.elements])
    t0 = time.perf_counter()
    n = 0
    dEn = []
    while n < obs.shape[0] and (n == 0 or time.perf_counter() - t0 < numpy_budget_s):
        zm, xm = obs[n]
        dEn.append(O.csr_wake(P.t + zm, xm, P.t, beam, P.formation_length, P.n_formation_length,
                              ip.xbins, ip.zbins, lat, hh, P.csr_scaling)[0])
        n += 1
    dtn = time.perf_counter() - t0
    res["numpy_1core"] = {"value": n * per_point / dtn / 1e9, "unit": "Gpts/s", "cores": 1,
                          "sample": f"oracle numpy restatement on {n} mesh points ({dtn:.1f} s)",
                          "max_rel_diff_vs_gpu_dE": rel_err(gpu_dE[:n], np.array(dEn))[0]}
    return res, dE


def profiled_traffic(name="wake_traffic.json"):
    """HBM bytes per launch (the wake's four, or one other kernel's) from the
    committed rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, each divided
    by its copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return int(d["fetch_kib_raw"] * 1024 / d["fetch_factor"] +
               d["write_kib_raw"] * 1024 / d["write_factor"])


def rel_err(a, ref):
    """(max, RMS) of |a - ref| over max|ref| on the entries where ref is
    finite, and whether the NaN positions of a and ref coincide (the
    reference's wake is NaN at mesh points past the lattice end)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    ok = np.isfinite(ref)
    if not ok.any():
        return float("nan"), float("nan"), bool(np.array_equal(np.isnan(a), np.isnan(ref)))
    d = a[ok] - ref[ok]
    m = np.abs(ref[ok]).max()
    return (float(np.abs(d).max() / m), float(np.sqrt(np.mean(d * d)) / m),
            bool(np.array_equal(np.isnan(a), np.isnan(ref))))


def compare_precisions(sim, reps=3):
    """FP64 and mixed wake on the simulation's current state (same history,
    same parameters): best-of-reps kernel time and the mixed form's error."""
    import numpy as np
    import torch
    from pydfcsr_amd import _lib, kernels
    H = sim.DF_tracker.descriptor()
    L = sim.lattice.device_tables(sim.device)
    P = sim.wake_params()
    pts = sim.integrand_points()
    res = {}
    outs = {}
    for name, prec in (("fp64", _lib.DFCSR_PREC_FP64), ("mixed", _lib.DFCSR_PREC_MIXED)):
        ts = []
        for _ in range(reps + 1):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            o = kernels.wake_mesh(H, L, P, prec, sim.device, sim.ws)
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        outs[name] = {k: v.cpu().numpy() for k, v in o.items()}
        t = min(ts[1:])
        res[name] = {"wake_ms": round(t, 3), "gpts": round(pts / (t * 1e-3) / 1e9, 3)}
    ref = outs["fp64"]
    for k in ("dE_dct", "x_kick"):
        mx, rms, same_nan = rel_err(outs["mixed"][k], ref[k])
        res["mixed"][f"max_rel_err_{k}"] = mx
        res["mixed"][f"rms_rel_err_{k}"] = rms
    res["mixed"]["nan_mesh_points"] = int(np.isnan(ref["dE_dct"]).sum())
    res["mixed"]["nan_positions_match"] = same_nan
    res["mixed_speedup"] = round(res["fp64"]["wake_ms"] / res["mixed"]["wake_ms"], 3)
    return res, outs


def serving_sample(dev, n, rate, steps_range, seconds, seed=0, overlap=True):
    """The serving mix (mix_requests: particles 1e5-1e7 log-uniform,
    deposition 64/128/256, wake mesh 32/64/128, single dipole rho 5-20 m or
    chicane at 0.02-0.08 rad, sigma 20-100 um, chirp -20..0 /m, history grid
    capped at 600 per side) with Poisson arrivals at `rate` per second, mixed
    precision, admission by HBM footprint, continuous batching; runs until
    every request finished or `seconds` of wall time.  Returns the server's
    report (throughput, step and request latency p50/p99, queue waits)."""
    from pydfcsr_amd.serving import Server, mix_requests
    reqs = mix_requests(n=n, rate=rate, seed=seed, steps_range=steps_range)
    srv = Server(device=dev, precision="mixed", use_graph=True, overlap=overlap)
    for r in reqs:
        srv.submit(r)

    def progress():
        return (f"serving: t={time.perf_counter() - t0:.1f}s ticks={srv.ticks} steps={srv.steps} "
                f"active={len(srv.active)} waiting={len(srv.waiting)} finished={len(srv.finished)}")

    t0 = time.perf_counter()
    last = t0
    start = srv.clock()
    while (srv.waiting or srv.active) and time.perf_counter() - t0 < seconds:
        now = srv.clock() - start
        if not srv.active and srv.waiting and srv.waiting[0].arrival > now:
            time.sleep(min(0.05, srv.waiting[0].arrival - now))
            continue
        srv.tick(now)
        if time.perf_counter() - last > 10:
            print(progress(), file=sys.stderr, flush=True)
            last = time.perf_counter()
    rep = srv.report(time.perf_counter() - t0)
    rep["requests"] = n
    rep["arrival_rate_per_s"] = rate
    rep["steps_range"] = list(steps_range)
    rep["overlap"] = overlap
    rep["history_grid_cap"] = 600
    srv.pool and srv.pool.shutdown()
    return rep


def serve_main(args, dev):
    """--workload serving: the serving mix alone (serving_sample)."""
    lo, hi = (int(v) for v in args.serve_steps.split(","))
    rep = serving_sample(dev, args.serve_requests, args.serve_rate, (lo, hi), args.serve_seconds,
                         overlap=not args.serve_serial)
    res = {"metric": METRIC, "value": rep["steps_per_s"], "unit": "sim-steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f32+f64", "data": "synthetic request mix",
           "config": {"workload": f"serving mix: {args.serve_requests} requests, Poisson "
                                  f"{args.serve_rate}/s, steps {lo}-{hi}, mixed precision, "
                                  f"history grid capped at 600 per side, "
                                  f"{args.serve_seconds:.0f} s wall budget"},
           "serving": rep}
    print(json.dumps(res), flush=True)


SG2D_GRID, SG2D_WINDOW, SG2D_ORDER = 8192, 7, 2
SG2D_BYTES_PER_POINT = 24    # 'both': one float64 read, two float64 writes


def sgolay_main(args, dev):
    """The 2D Savitzky-Golay filter (SGolay_filter.sgolay2d) as get_DF's
    commented alternative calls it for the gradients (derivative 'both',
    deposit.py:227, 232): one launch on an 8192 x 8192 float64 grid, window
    7, order 2; also the three calls get_DF would make per step at the
    chicane's 128 x 128 deposition grid (smoothing, 'both' on the density,
    'both' on vx)."""
    import numpy as np
    import torch
    from pydfcsr_amd import kernels
    n, w, order = SG2D_GRID, SG2D_WINDOW, SG2D_ORDER
    gen = torch.Generator(device=dev).manual_seed(0)
    z = torch.rand((n, n), dtype=torch.float64, device=dev, generator=gen)
    for _ in range(args.warmup if args.warmup is not None else 3):
        kernels.sgolay2d(z, w, order, "both")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)     # the stream the library launches on
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        kernels.sgolay2d(z, w, order, "both")
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    achieved = SG2D_BYTES_PER_POINT * n * n / (ms * 1e-3) / 1e9
    small = torch.rand((128, 128), dtype=torch.float64, device=dev, generator=gen)
    vx = torch.rand((128, 128), dtype=torch.float64, device=dev, generator=gen)
    use = []
    for _ in range(20):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        kernels.sgolay2d(small, w, order, None)
        kernels.sgolay2d(small, w, order, "both")
        kernels.sgolay2d(vx, w, order, "both")
        b.record(stream)
        torch.cuda.synchronize()
        use.append(a.elapsed_time(b) * 1e3)
    res = {"metric": "sgolay2d grid points/s", "value": round(n * n * args.steps / wall / 1e9, 4),
           "unit": "Gpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup if args.warmup is not None else 3,
           "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (uniform)",
           "config": {"workload": f"sgolay2d derivative 'both' on a {n} x {n} float64 grid, "
                                  f"window {w}, order {order}"},
           "kernel_ms": round(ms, 4),
           "get_df_calls_128x128_us": round(float(np.median(use[5:])), 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                        "traffic": profiled_traffic("sgolay2d_traffic.json"),
                        "kernel": "sgolay2d_kernel",
                        "bytes_per_point": SG2D_BYTES_PER_POINT}}
    if not args.no_cpu_baseline:
        from oracle import cpu_ref as O
        m = 2048
        zs = z[:m, :m].cpu().numpy()
        t = time.perf_counter()
        reps = 0
        while reps == 0 or time.perf_counter() - t < min(args.cpu_budget, 10.0):
            O.sgolay2d(zs, w, order, "both")
            reps += 1
        dt = (time.perf_counter() - t) / reps
        res["cpu_baseline"] = {"value": round(m * m / dt / 1e9, 5), "unit": "Gpts/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle sgolay2d (numpy pinv + scipy fftconvolve, as "
                                         f"the reference) on a {m} x {m} corner of the grid, "
                                         f"{reps} calls"}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed steps after the fill, before the timed ones")
    ap.add_argument("--fill", type=int, default=None,
                    help="steps run first to bring the history to its steady state (default: "
                         "450 large, 200 chicane, 90 dipole)")
    ap.add_argument("--workload", default="large",
                    choices=["chicane", "dipole", "large", "serving", "sgolay2d"])
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=90.0)
    ap.add_argument("--serve-steps", default="100,800", help="history length range (steps)")
    ap.add_argument("--serve-serial", action="store_true",
                    help="run the simulations' host steps one after another (no overlap)")
    ap.add_argument("--serve-sample-s", type=float, default=90.0,
                    help="wall budget of the serving sample the default run appends (0: none)")
    ap.add_argument("--serve-sample-requests", type=int, default=12)
    ap.add_argument("--serve-sample-rate", type=float, default=0.5)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device 3 times (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU mesh "
              "split is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pydfcsr_amd.csr import CSR2D

    if args.copy_calibration:
        src = torch.ones(args.copy_calibration // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return

    dev = torch.device("cuda:0")
    if args.workload == "serving":
        serve_main(args, dev)
        return
    if args.workload == "sgolay2d":
        sgolay_main(args, dev)
        return
    make_cfg, default_fill, description = WORKLOADS[args.workload]
    fill = default_fill if args.fill is None else args.fill
    cfg = make_cfg()
    t_init = time.perf_counter()
    # both precisions' rings: the precision comparison runs on the same state
    sim = CSR2D(config=cfg, device=dev, precision=args.precision, keep_fp32=True,
                keep_fp64=True)
    it = sim.iter_steps()
    for i in range(fill + args.warmup):
        next(it)
        if (i + 1) % 50 == 0 or i + 1 == fill:
            print(f"{'fill' if i < fill else 'warmup'} step {i + 1}/{fill + args.warmup}: "
                  f"s = {sim.beam.position:.4f} m, history {sim.DF_tracker.nt} slices on "
                  f"{sim.DF_tracker.Xi}x{sim.DF_tracker.Zi}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    s_start = float(sim.beam.position)
    slices_start = sim.DF_tracker.nt
    sim.timing = True
    sim.timing_events.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        next(it)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sim.timing = False
    ev = sim.timing_events
    csr_ms = [a.elapsed_time(d) for a, _, _, d, _, _ in ev]
    wake_ms = [b.elapsed_time(c) for _, b, c, _, _, _ in ev]
    pts = [p for *_, p, _ in ev]
    gpts = sum(pts) / (sum(wake_ms) * 1e-3) / 1e9
    wake_avg_ms = float(np.mean(wake_ms))
    tr = sim.DF_tracker
    # points the kernels evaluate on the final state (the same state the
    # precision comparison and the CPU legs use), per wake
    from pydfcsr_amd import _lib, kernels
    sim.get_CSR_mesh()
    H = tr.descriptor()
    Ld = sim.lattice.device_tables(sim.device)
    Pw = sim.wake_params()
    evaluated = {name: kernels.wake_points_evaluated(H, Ld, Pw, prec, sim.device, sim.ws)
                 for name, prec in (("fp64", _lib.DFCSR_PREC_FP64),
                                    ("mixed", _lib.DFCSR_PREC_MIXED))}
    # the points inside the history whose gathered density and gradients the
    # span table shows to be 0 (an integrand of exactly 0): skipped
    H_all = _lib.History.from_buffer_copy(H)
    H_all.zspan = None
    zero_skipped = kernels.wake_points_evaluated(H_all, Ld, Pw, _lib.DFCSR_PREC_FP64, sim.device,
                                                 sim.ws) - evaluated["fp64"]
    nominal = sim.integrand_points()
    frac_eval = evaluated["fp64"] / nominal
    prec_name = "fp64" if args.precision == "fp64" else "mixed"
    achieved = FLOPS_PER_POINT * evaluated[prec_name] / (wake_avg_ms * 1e-3) / 1e12
    if args.precision == "fp64":
        peak, rkernel = FP64_PEAK_TFLOPS, "FP64 wake"
    else:   # the mixed form's ops at both peaks: the equivalent single-peak rate
        t_ideal = evaluated["mixed"] * ((FLOPS_PER_POINT - MIXED_FP64_OPS) / FP32_PEAK_TFLOPS +
                                        MIXED_FP64_OPS / FP64_PEAK_TFLOPS) / 1e12
        peak = FLOPS_PER_POINT * evaluated["mixed"] / t_ideal / 1e12
        rkernel = "mixed-precision wake"
    traffic = profiled_traffic(f"wake_traffic_{args.workload}.json") \
        if args.precision == "fp64" else None
    res = {
        "metric": METRIC,
        "value": round(gpts, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": f"synthetic (seeded beam, {args.workload} lattice)",
        "config": {"workload": description,
                   "precision": args.precision,
                   "fill_steps": fill,
                   "steps_timed": f"{ev[0][5] if ev else 0} mesh points x "
                                  f"{(pts[0] // ev[0][5]) if ev else 0} nominal points per step",
                   "s_start_m": round(s_start, 6),
                   "slices_at_start": slices_start},
        "csr_step_ms": round(float(np.mean(csr_ms)), 3),
        "wake_ms": round(wake_avg_ms, 3),
        "points_nominal_per_wake": nominal,
        "points_evaluated_per_wake": evaluated,
        "points_zero_skipped_per_wake": zero_skipped,
        "history": {"slices": tr.nt, "Xi": tr.Xi, "Zi": tr.Zi,
                    "bytes": int(tr.footprint_bytes()), "regrids": tr.n_regrid,
                    "cell_major": [k for k in ("cells", "cells32")
                                   if getattr(tr, k) is not None]},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                     "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     # bytes per wake from the committed counter runs of this state
                     "traffic": traffic,
                     "kernel": f"{rkernel} (wake_observer_kernel, wake_far_kernel for region 0, "
                               "wake_phase_kernel for the near regions, wake_finalize_kernel; "
                               "one stream, timed together)",
                     "flops_per_point": FLOPS_PER_POINT,
                     "points": "evaluated",
                     "nominal_equivalent_frac": round(achieved / frac_eval / peak, 4)},
        "init_s": round(t0 - t_init, 2),
    }
    print("precision comparison: fp64 and mixed wake on the final state", file=sys.stderr,
          flush=True)
    res["precision_compare"], outs = compare_precisions(sim)
    pc = res["precision_compare"]
    pc["fp64"]["tflops_evaluated"] = round(
        FLOPS_PER_POINT * evaluated["fp64"] / (pc["fp64"]["wake_ms"] * 1e-3) / 1e12, 3)
    t_ideal = evaluated["mixed"] * ((FLOPS_PER_POINT - MIXED_FP64_OPS) / (FP32_PEAK_TFLOPS * 1e12)
                                    + MIXED_FP64_OPS / (FP64_PEAK_TFLOPS * 1e12))
    pc["mixed"]["roofline"] = {
        "fp32_ops_per_point": FLOPS_PER_POINT - MIXED_FP64_OPS,
        "fp64_ops_per_point": MIXED_FP64_OPS,
        "fp32_tflops": round((FLOPS_PER_POINT - MIXED_FP64_OPS) * evaluated["mixed"] /
                             (pc["mixed"]["wake_ms"] * 1e-3) / 1e12, 3),
        "fp64_tflops": round(MIXED_FP64_OPS * evaluated["mixed"] /
                             (pc["mixed"]["wake_ms"] * 1e-3) / 1e12, 3),
        "frac": round(t_ideal / (pc["mixed"]["wake_ms"] * 1e-3), 4)}
    if not args.no_cpu_baseline:
        print("cpu baseline: oracle C/OpenMP and numpy integrals on the same history",
              file=sys.stderr, flush=True)
        res["cpu_baseline"], dE_cpu = cpu_baseline(sim, outs["fp64"], budget_s=args.cpu_budget)
        # the mixed form against the FP64 CPU reference on the same sample
        k = dE_cpu.shape[0]
        mx, rms, _ = rel_err(outs["mixed"]["dE_dct"][:k], dE_cpu)
        pc["mixed"]["max_rel_err_dE_dct_vs_cpu"] = mx
        pc["mixed"]["rms_rel_err_dE_dct_vs_cpu"] = rms
    if args.serve_sample_s > 0:
        # the serving mix, bounded: a seeded sample of the request stream at a
        # rate the server keeps up with (queue waits stay near zero)
        del sim, outs
        torch.cuda.empty_cache()
        lo, hi = (int(v) for v in args.serve_steps.split(","))
        print(f"serving sample: {args.serve_sample_requests} requests at "
              f"{args.serve_sample_rate}/s", file=sys.stderr, flush=True)
        res["serving"] = serving_sample(dev, args.serve_sample_requests, args.serve_sample_rate,
                                        (lo, hi), args.serve_sample_s)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
