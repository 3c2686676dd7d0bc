"""Full-size parity runs of the single-dipole configuration (plain FP64 and its
mixed-precision form): 2e5 particles, 5 GeV, rho = 10 m, L = 0.5 m, 100
steps, deposition 64 x 64 SG 5/2, wake mesh 32 x 32, integration 200 x 200.

The driver runs all 100 steps.  The oracle keeps its own history
(deposit.py:145-426, numpy) fed with the particles the driver tracked at each
step; at three steps the oracle's C/OpenMP restatement of the integral
(CSR.py:397-602) evaluates every mesh point on the oracle's history and the
driver's FP64 and mixed-precision wakes are compared with it.  Tolerances:
FP64 1e-9 of max|W|; mixed 1e-4 of max|W| (the stated precision target).
At the last step some mesh points lie past the lattice end.  Beyond the
table's last interval the reference's lookups are 0 and its integrand 0/0
(NaN, matched position by position); within that interval (L < s < L + ds)
observer and nearby sources look up the same clamped table row, R reduces
to |x - x'| and the 1/R^3 terms make the reference formula itself
ill-conditioned: there the FP64 bound is 1e-8.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _max_rms(a, b, sel=None):
    """Max and RMS of |a - b| relative to max|b| over the finite entries
    (restricted to sel for the max and RMS); NaN (the reference's 0/0 at mesh
    points past the lattice end) must coincide."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nan = np.isnan(b)
    assert np.array_equal(np.isnan(a), nan), \
        f"NaN positions differ: mine {np.flatnonzero(np.isnan(a))[:8].tolist()} " \
        f"reference {np.flatnonzero(nan)[:8].tolist()}"
    m = np.abs(b[~nan]).max()
    ok = ~nan if sel is None else (~nan & sel)
    if not ok.any():
        return 0.0, 0.0
    d = a[ok] - b[ok]
    return float(np.abs(d).max() / m), float(np.sqrt(np.mean(d * d)) / m)


def test_dipole_full_run_fp64_and_mixed_vs_oracle():
    import torch
    from bench import dipole_config
    from oracle import cpu_ref as O
    from oracle import csr_ref_c as C
    from pydfcsr_amd import _lib, kernels
    from pydfcsr_amd.csr import CSR2D

    from tests.gpu_util import LoopReplay

    cfg = dipole_config()
    dep, ip, cp = cfg["particle_deposition"], cfg["CSR_integration"], cfg["CSR_computation"]
    sim = CSR2D(config=cfg, device=DEV, precision="fp64", keep_fp32=True)
    loop = LoopReplay(sim, cfg["input_lattice"]["_config"])
    ora = O.DFTracker(xbins=dep["xbins"], zbins=dep["zbins"], xlim=dep["xlim"],
                      zlim=dep["zlim"], filter_order=dep["filter_order"],
                      filter_window=dep["filter_window"],
                      velocity_threhold=dep["velocity_threhold"])
    b = sim.beam
    ora.get_DF(_np(b.x), _np(b.z), _np(b.px), float(b.position))
    ora.append_DF()
    ora.append_interpolant(float("inf"), ip["n_formation_length"])
    lat = O.LatticeTables([dict(v) for k, v in cfg["input_lattice"]["_config"].items()
                           if k != "step_size"])
    check = {25, 60, 100}
    report = {}
    for kind, step in sim.iter_phases(external_wake=True):
        if kind != "wake":
            continue
        x, z, px = _np(b.x), _np(b.z), _np(b.px)
        t = float(b.position)
        fl = loop.step(step)["formation_length"]
        assert abs(fl - sim.formation_length) <= 1e-13 * fl, (step, fl, sim.formation_length)
        ora.get_DF(x, z, px, t)
        ora.append_DF()
        ora.append_interpolant(fl, ip["n_formation_length"])
        H, L, P, out = sim.prepare_wake()
        kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_FP64, sim.device, sim.ws, out=out,
                          want_mesh=False)
        if step not in check:
            continue
        print(f"step {step}: s = {t:.3f} m, {sim.DF_tracker.nt} slices", flush=True)
        mixed = kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_MIXED, sim.device, sim.ws,
                                  want_mesh=False)
        torch.cuda.synchronize()
        ora.build_interpolant()
        keep = C._Keep()
        hist = C.hist_from_tracker(ora, keep)
        latc = C.lat_from_tables(lat, keep)
        mesh = O.csr_mesh(x, z, cp["xbins"], cp["zbins"], cp["xlim"], cp["zlim"])
        m = O.beam_moments(x, z)
        beam = dict(sigma_x=m["sigma_x"], sigma_z=m["sigma_z"], slope=m["slope"],
                    mean_x=m["mean_x"])
        obs = np.stack([mesh["zmesh"], mesh["xmesh"]], axis=1)
        dE, xk = C.wake(hist, latc, t, beam, fl, ip["n_formation_length"],
                        ip["xbins"], ip["zbins"], sim.CSR_scaling, obs, C.threads())
        inside = t + mesh["zmesh"] <= lat.s[-1]
        for name, res in (("fp64", out), ("mixed", mixed)):
            for k, ref in (("dE_dct", dE), ("x_kick", xk)):
                report[(step, name, k)] = _max_rms(_np(res[k]), ref)
                report[(step, name, k, "s<=L")] = _max_rms(_np(res[k]), ref, inside)
        print({k: v for k, v in report.items() if k[0] == step}, flush=True)
        for k in ("dE_dct", "x_kick"):
            assert report[(step, "fp64", k, "s<=L")][0] <= 1e-9, (step, k, report)
            assert report[(step, "fp64", k)][0] <= 1e-8, (step, k, report)
            assert report[(step, "mixed", k)][0] <= 1e-4, (step, k, report)
    assert check <= {k[0] for k in report}


def _deep_run(cfg, steps, rows_seed, wake_tols, regrid_slices=3, check_get_df=True):
    """Runs the driver (FP64, float32 rings kept too) for `steps` steps with
    the formation length checked against the oracle's run loop, and at the
    last step, on the state its wake reads:
      * the device's slices re-gridded independently: sampled raw slices
        (oldest, middle, newest of the window) through scipy's
        RegularGridInterpolator as DF_interp does (deposit.py:296-309), fill 0
        and mean(vx_x) for vx_x, against the float64 ring (1e-12 of max);
      * the newest raw slice against the oracle's get_DF (deposit.py:145-245)
        on the particles of that step (fields 1e-9 of max, grids exact);
      * FP64 and mixed wakes at 64 seeded mesh points against the oracle's
        C/OpenMP integral on the device's history with the oracle's
        formation length (wake_tols = (fp64, mixed) of max|W|)."""
    import torch
    from scipy.interpolate import RegularGridInterpolator
    from oracle import cpu_ref as O
    from oracle import csr_ref_c as C
    from pydfcsr_amd import _lib, kernels
    from pydfcsr_amd.csr import CSR2D
    from tests.gpu_util import LoopReplay

    sim = CSR2D(config=cfg, device=DEV, precision="fp64", keep_fp32=True)
    loop = LoopReplay(sim, cfg["input_lattice"]["_config"])
    report = {}
    for kind, step in sim.iter_phases(external_wake=True):
        if kind == "done":
            if step == steps:
                break
            continue
        H, L, P, out = sim.prepare_wake()
        kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_FP64, sim.device, sim.ws, out=out,
                          want_mesh=False)
        if step < steps:
            continue
        torch.cuda.synchronize()
        tr = sim.DF_tracker
        rec = loop.step(step)
        assert rec["position"] == P.t, (rec, P.t)
        fl = rec["formation_length"]
        assert abs(fl - P.formation_length) <= 1e-13 * fl, (fl, P.formation_length)
        print(f"step {step}: s = {P.t:.4f} m, {tr.nt} slices on {tr.Xi}x{tr.Zi}, "
              f"{tr.n_regrid} full re-grids", flush=True)
        # -- independent re-grid of sampled slices
        hf = tr.history_fields()
        X = np.linspace(tr.xi[0], tr.xi[1], tr.Xi)
        Z = np.linspace(tr.zi[0], tr.zi[1], tr.Zi)
        XX, ZZ = np.meshgrid(X, Z, indexing="ij")
        picks = sorted({0, tr.nt // 2, tr.nt - 1})[:regrid_slices]
        worst = 0.0
        for i in picks:
            m = tr.log[i]
            raw = [_np(f) for f in tr.raw_fields(i)]
            xg, zg = np.linspace(m.x0, m.x1, m.nbx), np.linspace(m.z0, m.z1, m.nbz)
            for f, name in enumerate(O.FIELDS):
                fill = np.mean(raw[f]) if name == "vx_x" else 0.0
                ref = RegularGridInterpolator((xg, zg), raw[f], method="linear",
                                              fill_value=fill, bounds_error=False)((XX, ZZ))
                mine = _np(hf[name][i])
                den = np.abs(ref).max()
                e = float(np.abs(mine - ref).max() / den) if den > 0 else float(
                    np.abs(mine - ref).max())
                worst = max(worst, e)
                assert e <= 1e-12, (i, name, e)
        report["regrid"] = worst
        del hf
        # -- the newest slice against the oracle's get_DF on this step's particles
        if check_get_df:
            b = sim.beam
            dep = cfg["particle_deposition"]
            ref = O.get_DF(_np(b.x), _np(b.z), _np(b.px), P.t, **dep)
            m = tr.log[-1]
            assert (m.nbx, m.nbz) == ref["nbins"] and m.window == ref["window"]
            # grid ends from the device's moments and numpy's: equal to rounding
            assert np.allclose((m.x0, m.x1, m.z0, m.z1), (*ref["x_range"], *ref["z_range"]),
                               rtol=1e-12, atol=0)
            raw = [_np(f) for f in tr.raw_fields(tr.nt - 1)]
            e = max(_max_rms(raw[f], ref[name])[0] for f, name in enumerate(O.FIELDS))
            report["get_DF"] = e
            assert e <= 1e-9, e
        # -- the wakes at sampled mesh points
        mixed = kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_MIXED, sim.device, sim.ws,
                                  want_mesh=False)
        fp64 = kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_FP64, sim.device, sim.ws)
        torch.cuda.synchronize()
        M = P.mesh_xbins * P.mesh_zbins
        rows = np.sort(np.random.default_rng(rows_seed).choice(M, 64, replace=False))
        keep = C._Keep()
        hist = C.hist_from_records(tr.hist.cpu().numpy(), tr.head, tr.nt,
                                   (float(tr.min_x), float(tr.min_y), float(tr.min_z),
                                    float(tr.delta_x), float(tr.delta_y), float(tr.delta_z)),
                                   keep)
        latc = C.lat_from_tables(sim.lattice, keep)
        bm = sim.beam
        beam = dict(sigma_x=bm._sigma_x, sigma_z=bm._sigma_z, slope=bm._slope,
                    mean_x=bm._mean_x)
        obs = np.stack([_np(fp64["zmesh"])[rows], _np(fp64["xmesh"])[rows]], axis=1)
        ip = cfg["CSR_integration"]
        dE, xk = C.wake(hist, latc, P.t, beam, fl, P.n_formation_length, ip["xbins"],
                        ip["zbins"], P.csr_scaling, obs, C.threads())
        # mesh points past the lattice end: NaN where the reference's lookups
        # are all 0, and ill-conditioned in the table's last interval (as in
        # the single-dipole test); the strict bound holds for s <= L
        inside = P.t + obs[:, 0] <= sim.lattice.lattice_length
        for name, res in (("fp64", fp64), ("mixed", mixed), ("fp64_driver", out)):
            for k, ref in (("dE_dct", dE), ("x_kick", xk)):
                report[(name, k)] = _max_rms(_np(res[k])[rows], ref)
                report[(name, k, "s<=L")] = _max_rms(_np(res[k])[rows], ref, inside)
        print(report, flush=True)
        for k in ("dE_dct", "x_kick"):
            for name in ("fp64", "fp64_driver"):
                assert report[(name, k, "s<=L")][0] <= wake_tols[0], report
                assert report[(name, k)][0] <= max(wake_tols[0], 1e-8), report
            assert report[("mixed", k)][0] <= wake_tols[1], report
        return sim, report
    raise AssertionError(f"the run ended before step {steps}")


def test_large_run_to_steady_state_vs_oracle():
    """Full-size large configuration (1e7 particles, deposition 256 x 256 SG
    9/3, wake mesh 128 x 128, 200 x 200 integration) run to its steady state
    (455 steps: 430 slices on 500 x 500, the window dropping slices): every
    check of _deep_run; the last slice's density integrates to 1."""
    import torch
    from bench import large_config

    sim, report = _deep_run(large_config(), 455, 5, (1e-9, 1e-4))
    tr = sim.DF_tracker
    assert tr.nt >= 400 and tr.log[0].t > 0   # slices have been dropped
    c = tr.log[-1]
    d = _np(tr.raw_fields(tr.nt - 1)[0])
    xg = np.linspace(c.x0, c.x1, c.nbx)
    zg = np.linspace(c.z0, c.z1, c.nbz)
    trapz = getattr(np, "trapezoid", None) or np.trapz
    integral = trapz(trapz(d, xg, axis=0), zg)
    assert abs(integral - 1.0) <= 1e-12, integral
    # with and without the span table (the near regions' lookup skip):
    # bit-identical wakes, fewer points evaluated
    from pydfcsr_amd import _lib, kernels
    H = tr.descriptor()
    H0 = _lib.History.from_buffer_copy(H)
    H0.zspan = None
    P = sim.wake_params()
    L = sim.lattice.device_tables(sim.device)
    for prec in (_lib.DFCSR_PREC_FP64, _lib.DFCSR_PREC_MIXED):
        a = kernels.wake_mesh(H, L, P, prec, sim.device, sim.ws, want_mesh=False)
        b = kernels.wake_mesh(H0, L, P, prec, sim.device, sim.ws, want_mesh=False)
        for k in ("dE_dct", "x_kick"):
            same = (a[k] == b[k]) | (torch.isnan(a[k]) & torch.isnan(b[k]))
            assert bool(same.all()), (prec, k)
        ev = [kernels.wake_points_evaluated(h, L, P, prec, sim.device, sim.ws) for h in (H, H0)]
        assert ev[0] < ev[1], (prec, ev)
    torch.cuda.synchronize()


def test_chicane_full_run_vs_oracle():
    """The chicane configuration (1e6 particles, 4 bends at 0.05 rad,
    deposition 128 x 128 SG 7/2, wake mesh 64 x 64, 200 x 200 integration,
    chirp branch mid-chicane) run through all its steps (399: the reference's
    step schedule, lattice.py:152-173, loses one position to rounding at an
    element boundary): every check of _deep_run at the last step (the beam
    at the lattice end: mesh points past it give the reference's NaN, which
    must coincide)."""
    from bench import chicane_config
    from pydfcsr_amd.lattice import Lattice
    cfg = chicane_config()
    last = int(Lattice(cfg["input_lattice"]).steps_per_element.sum())
    assert last == 399
    sim, report = _deep_run(cfg, last, 7, (1e-9, 1e-4))
    assert sim.DF_tracker.nt > 150


def test_chicane_mid_run_chirp_branch_vs_oracle():
    """The chicane at step 200 (192 slices on the full interp grid, chirp
    branch: four regions): FP64 and mixed wakes against the oracle as in
    _deep_run, and the mixed wake without the float64 ring (vx at the
    observer from float32) within 1e-4 as well; the wakes with and without
    the span table bit-identical."""
    import torch
    from bench import chicane_config
    from pydfcsr_amd import _lib, kernels
    sim, report = _deep_run(chicane_config(), 200, 7, (1e-9, 1e-4), regrid_slices=2,
                            check_get_df=False)
    H = sim.DF_tracker.descriptor()
    P = sim.wake_params()
    assert abs(P.slope0) > 1   # the chirp branch (four regions)
    L = sim.lattice.device_tables(sim.device)
    H32 = _lib.History.from_buffer_copy(H)
    H32.data = None
    ref = kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_MIXED, sim.device, sim.ws, want_mesh=False)
    mixed32 = kernels.wake_mesh(H32, L, P, _lib.DFCSR_PREC_MIXED, sim.device, sim.ws,
                                want_mesh=False)
    torch.cuda.synchronize()
    for k in ("dE_dct", "x_kick"):
        assert _max_rms(_np(mixed32[k]), _np(ref[k]))[0] <= 1e-4
    # with and without the span table (on the chirp branch region 0 skips on it as well
    # as the near regions): bit-identical wakes, fewer points evaluated
    H0 = _lib.History.from_buffer_copy(H)
    H0.zspan = None
    for prec in (_lib.DFCSR_PREC_FP64, _lib.DFCSR_PREC_MIXED):
        a = kernels.wake_mesh(H, L, P, prec, sim.device, sim.ws, want_mesh=False)
        b = kernels.wake_mesh(H0, L, P, prec, sim.device, sim.ws, want_mesh=False)
        for k in ("dE_dct", "x_kick"):
            same = (a[k] == b[k]) | (torch.isnan(a[k]) & torch.isnan(b[k]))
            assert bool(same.all()), (prec, k)
        ev = [kernels.wake_points_evaluated(h, L, P, prec, sim.device, sim.ws) for h in (H, H0)]
        assert ev[0] < ev[1], (prec, ev)
