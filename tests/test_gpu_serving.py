"""GPU tests of the serving loop: a simulation's wakes are identical alone
and batched with others; graph replays of the batched launch equal eager
launches on every replay."""
from collections import OrderedDict

import numpy as np
import pytest
import torch

from pydfcsr_amd.serving import BatchedWake, Request, Server

pytestmark = pytest.mark.gpu


def _cfg(n, nb, mesh, Nxz, kind, seed):
    lat = OrderedDict()
    if kind == "dipole":
        lat["step_size"] = 0.5 / 10
        lat["B1"] = dict(type="dipole", L=0.5, angle=0.05, nsep=1)
    else:
        lat["step_size"] = 6.2 / 60
        lat["B1"] = dict(type="dipole", L=0.3, angle=0.05, E1=0.0, E2=0.05, nsep=1)
        lat["D1"] = dict(type="drift", L=2.0, nsep=1)
        lat["B2"] = dict(type="dipole", L=0.3, angle=-0.05, E1=-0.05, E2=0.0, nsep=1)
        lat["D2"] = dict(type="drift", L=1.0, nsep=1)
        lat["B3"] = dict(type="dipole", L=0.3, angle=-0.05, E1=0.0, E2=-0.05, nsep=1)
        lat["D3"] = dict(type="drift", L=2.0, nsep=1)
        lat["B4"] = dict(type="dipole", L=0.3, angle=0.05, E1=0.05, E2=0.0, nsep=1)
    return {
        "input_beam": dict(style="gaussian", n_particle=n, charge=1e-9, energy=1e9,
                           sigma_x=40e-6, sigma_z=60e-6, chirp=-8.0, sigma_delta=1e-4, seed=seed),
        "input_lattice": dict(lattice_input_file="<t>", _config=lat),
        "particle_deposition": dict(xbins=nb, zbins=nb, filter_order=2, filter_window=5,
                                    velocity_threhold=1000, upper_limit=120),
        "CSR_integration": dict(n_formation_length=4, zbins=Nxz, xbins=Nxz),
        "CSR_computation": dict(xbins=mesh[0], zbins=mesh[1], xlim=4, zlim=4),
    }


REQS = [(_cfg(20000, 64, (6, 8), 24, "dipole", 1), 8),
        (_cfg(30000, 64, (4, 6), 20, "chicane", 2), 10),
        (_cfg(15000, 64, (8, 8), 16, "dipole", 3), 6)]


def _serve(indices, arrivals, precision="mixed", use_graph=True):
    srv = Server(device="cuda", hbm_budget=8 << 30, max_active=8, precision=precision,
                 use_graph=use_graph, record_wakes=True)
    for i, a in zip(indices, arrivals):
        cfg, steps = REQS[i]
        srv.submit(Request(i, cfg, a, steps, precision, footprint=1 << 20))
    rep = srv.run(simulated_clock=True, tick_dt=0.01)
    return srv, rep


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_results_same_alone_and_in_mix(precision):
    srv, rep = _serve([0, 1, 2], [0.0, 0.015, 0.035], precision)
    assert rep["finished"] == 3 and rep["sim_steps"] == 8 + 10 + 6
    assert rep["mean_active"] > 1.5   # the simulations really overlapped
    for i in range(3):
        alone, _ = _serve([i], [0.0], precision, use_graph=False)
        a = alone.finished[i].wake_results
        m = srv.finished[i].wake_results
        assert len(a) == len(m) > 0
        for (dEa, xka), (dEm, xkm) in zip(a, m):
            assert torch.equal(dEa, dEm) and torch.equal(xka, xkm)


def test_batched_matches_single_kernel():
    """The batched launch equals the per-simulation kernel on the same state."""
    from pydfcsr_amd import kernels
    srv, _ = _serve([0], [0.0], "fp64")
    sim = srv.finished[0].sim
    H, L, P, out = sim.prepare_wake()
    BatchedWake("cuda", "fp64", use_graph=False).run([(H, L, P, out)])
    ref = kernels.wake_mesh(H, L, P, 0, sim.device, sim.ws)
    torch.cuda.synchronize()
    d = (out["dE_dct"] - ref["dE_dct"]).abs().max().item()
    assert d <= 1e-13 * ref["dE_dct"].abs().max().item()


def test_graph_replay_matches_eager():
    srv, _ = _serve([0, 1], [0.0, 0.0], "mixed")
    sims = [srv.finished[i].sim for i in (0, 1)]
    items = [s.prepare_wake() for s in sims]
    BatchedWake("cuda", "mixed", use_graph=False).run(items)
    torch.cuda.synchronize()
    ref = [(it[3]["dE_dct"].clone(), it[3]["x_kick"].clone()) for it in items]
    bw = BatchedWake("cuda", "mixed", use_graph=True)
    for rep in range(4):          # eager, capture + replay, replay, replay
        for it in items:
            it[3]["dE_dct"].fill_(np.nan)
            it[3]["x_kick"].fill_(np.nan)
        bw.run(items)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum().item()   # eager work between replays
        assert np.isfinite(junk)
        for it, (dE, xk) in zip(items, ref):
            assert torch.equal(it[3]["dE_dct"], dE) and torch.equal(it[3]["x_kick"], xk), rep
    assert bw.replays == 3
