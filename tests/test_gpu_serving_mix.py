"""The serving mix at its extremes, under the serving loop, on the GPU.

Four requests drawn from the seeded request stream (mix_requests, seed 0:
beam sizes, chirp, beam seeds) with their sizes set to the ranges' ends:
  * a 4-bend chicane at 0.08 rad, 1e7 particles, deposition 256 x 256, wake
    mesh 128 x 128, 800 steps;
  * a single dipole (rho 5 m), 1e5 particles, 64 x 64, 32 x 32, 100 steps;
  * a chicane at 0.02 rad, 1e6 particles, 128 x 128, 64 x 64, 300 steps;
  * a dipole (rho 20 m), 3e6 particles, 256 x 256, 128 x 128, 200 steps;
all mixed precision at 200 x 200 integration with the history grid capped at
600 per side (the stream's upper_limit), admitted on their computed HBM
footprints, continuous batching with the batched wake in HIP graphs.

Checked: every request finishes; each one's wakes at every step are the same
bit for bit alone (a server holding only it) and in the mix; each one's last
wake, recomputed from its recorded inputs, equals the recorded one and
agrees with the oracle's C/OpenMP FP64 integral on the same (float32) history
at 32 seeded mesh points within 1e-4 of max|W|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _requests():
    from pydfcsr_amd.serving import make_request, mix_requests
    base = mix_requests(n=4, rate=4.0, seed=0)
    ext = [("chicane", 0.08, 10_000_000, 256, 128, 800),
           ("dipole", 5.0, 100_000, 64, 32, 100),
           ("chicane", 0.02, 1_000_000, 128, 64, 300),
           ("dipole", 20.0, 3_000_000, 256, 128, 200)]
    out = []
    for r, (kind, strength, n, nb, mesh, steps) in zip(base, ext):
        b = r.config["input_beam"]
        out.append(make_request(r.rid, 0.0, steps, n, nb, mesh, kind, strength, b["sigma_x"],
                                b["sigma_z"], b["chirp"], b["seed"]))
    return out


class _HostHistory:
    def __init__(self, tr):
        self.data = {k: v.cpu().numpy() for k, v in tr.history_fields().items()}
        self._a = dict(min_t=float(tr.min_x), min_x=float(tr.min_y), min_z=float(tr.min_z),
                       dt=float(tr.delta_x), dx=float(tr.delta_y), dz=float(tr.delta_z))

    def axes(self):
        return self._a


def _same(a, b):
    """Bit-equal, NaN where the other is NaN (mesh points past the lattice
    end give the reference's 0/0)."""
    import torch
    na, nb = torch.isnan(a), torch.isnan(b)
    return bool(torch.equal(na, nb) and torch.equal(a[~na], b[~nb]))


def _serve(reqs):
    from pydfcsr_amd.serving import Request, Server
    srv = Server(device=DEV, precision="mixed", use_graph=True, record_wakes=True)
    for r in reqs:
        assert srv.submit(Request(r.rid, r.config, 0.0, r.steps, r.precision)), r.rid
    rep = srv.run(simulated_clock=True)
    return srv, rep


def test_serving_mix_extremes_alone_and_together_vs_oracle():
    import torch
    from oracle import csr_ref_c as C
    from pydfcsr_amd import _lib, kernels

    reqs = _requests()
    mix, rep = _serve(reqs)
    print("mix:", {k: rep[k] for k in ("sim_steps", "finished", "failed", "mean_active",
                                       "graph_replays")}, flush=True)
    assert rep["finished"] == 4 and rep["failed"] == 0 and rep["rejected"] == 0
    assert rep["mean_active"] > 1 and rep["graph_replays"] > 0
    from pydfcsr_amd.lattice import Lattice
    # a request runs to its step count or to its lattice's end, whichever is
    # first (the reference's schedule can hold one step less, lattice.py:152-173)
    n_steps = {r.rid: min(r.steps, int(Lattice(r.config["input_lattice"]).steps_per_element.sum()))
               for r in reqs}
    for r in reqs:
        slot = mix.finished[r.rid]
        assert len(slot.wake_results) == n_steps[r.rid] >= r.steps - 1
        # the last wake again, from its recorded inputs, and the oracle on it
        H, L, P = slot.last_wake
        sim = slot.sim
        again = kernels.wake_mesh(H, L, P, _lib.DFCSR_PREC_MIXED, sim.device, sim.ws)
        dE_last, xk_last = slot.wake_results[-1]
        for x, y in ((again["dE_dct"], dE_last.reshape(-1)), (again["x_kick"], xk_last.reshape(-1))):
            assert torch.equal(torch.isnan(x), torch.isnan(y))
            ok = ~torch.isnan(y)
            assert float((x[ok] - y[ok]).abs().max() / y[ok].abs().max()) <= 1e-12
        tr = sim.DF_tracker
        keep = C._Keep()
        hist = C.hist_from_tracker(_HostHistory(tr), keep)
        lat = C.lat_from_tables(sim.lattice, keep)
        M = P.mesh_xbins * P.mesh_zbins
        rows = np.sort(np.random.default_rng(r.rid).choice(M, 32, replace=False))
        xm, zm = again["xmesh"].cpu().numpy(), again["zmesh"].cpu().numpy()
        obs = np.stack([zm[rows], xm[rows]], axis=1)
        beam = dict(sigma_x=P.sigma_x, sigma_z=P.sigma_z, slope=np.array([P.slope0, P.slope1]),
                    mean_x=P.mean_x)
        dE, xk = C.wake(hist, lat, P.t, beam, P.formation_length, P.n_formation_length, P.Nx,
                        P.Nz, P.csr_scaling, obs, C.threads())
        for mine, ref in ((again["dE_dct"], dE), (again["x_kick"], xk)):
            a = mine.cpu().numpy()[rows]
            assert np.array_equal(np.isnan(a), np.isnan(ref))
            ok = ~np.isnan(ref)
            err = np.abs(a[ok] - ref[ok]).max() / np.abs(ref[ok]).max()
            print(f"request {r.rid}: {r.steps} steps, {tr.nt} slices on {tr.Xi}x{tr.Zi}, "
                  f"last mixed wake vs oracle {err:.2e}", flush=True)
            assert err <= 1e-4, (r.rid, err)
    # alone: each request in a server of its own, wakes bit for bit the mix's
    for r in reqs[1:]:
        alone, rep1 = _serve([r])
        assert rep1["finished"] == 1
        a, m = alone.finished[r.rid].wake_results, mix.finished[r.rid].wake_results
        assert len(a) == len(m) == n_steps[r.rid]
        for k, ((dA, xA), (dM, xM)) in enumerate(zip(a, m)):
            assert _same(dA, dM) and _same(xA, xM), (r.rid, k)
        del alone
        torch.cuda.empty_cache()
