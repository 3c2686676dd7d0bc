"""CPU tests of the serving loop's scheduling (admission by HBM footprint,
slots, continuous batching, retirement) with stub simulations."""
import pytest

from pydfcsr_amd.serving import Request, Server, mix_requests, request_footprint


class StubSim:
    def __init__(self, req, log):
        self.req = req
        self.log = log

    def iter_phases(self, external_wake=True):
        for s in range(1, self.req.steps + 1):
            yield ("wake", s)
            yield ("done", s)

    def prepare_wake(self):
        return (None, None, None, {})

    def external_wake(self, item):
        self.log.append(self.req.rid)

    def integrand_points(self):
        return 1000


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def _server(budget, max_active, log):
    return Server(device="cpu", hbm_budget=budget, max_active=max_active,
                  sim_factory=lambda r: StubSim(r, log), clock=Clock())


def test_admission_respects_budget_and_slots():
    log = []
    srv = _server(budget=100, max_active=2, log=log)
    for rid, (fp, steps) in enumerate([(60, 3), (30, 2), (50, 2), (10, 1)]):
        assert srv.submit(Request(rid, {}, arrival=0.0, steps=steps, footprint=fp))
    srv.tick(0.0)
    # 60 + 30 fit, 50 would exceed the budget; slots = 2
    assert [s.req.rid for s in srv.active] == [0, 1] and srv.used == 90
    srv.tick(0.0)   # request 1 finishes (2 steps)
    assert [s.req.rid for s in srv.active] == [0] and srv.used == 60
    srv.tick(0.0)   # FIFO: request 2 (50) does not fit beside 60 -> waits, 3 waits behind it
    assert [s.req.rid for s in srv.active] == [2] or srv.used <= 100
    rep = srv.run(simulated_clock=True)
    assert rep["finished"] == 4 and rep["still_waiting"] == 0 and srv.used == 0
    assert rep["sim_steps"] == 3 + 2 + 2 + 1
    # one batched wake per tick covers every active simulation
    assert sorted(set(log)) == [0, 1, 2, 3]


def test_requests_larger_than_budget_are_rejected():
    srv = _server(budget=100, max_active=4, log=[])
    assert not srv.submit(Request(0, {}, 0.0, 2, footprint=101))
    assert srv.rejected == [0]


def test_arrivals_wait_for_their_time():
    log = []
    srv = _server(budget=1000, max_active=8, log=log)
    srv.submit(Request(0, {}, arrival=0.0, steps=2, footprint=10))
    srv.submit(Request(1, {}, arrival=5.0, steps=2, footprint=10))
    srv.tick(0.0)
    assert [s.req.rid for s in srv.active] == [0]
    rep = srv.run(simulated_clock=True)
    assert rep["finished"] == 2


def test_mix_requests_ranges():
    reqs = mix_requests(n=64, rate=2.0, seed=3)
    assert len(reqs) == 64
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr)
    for r in reqs:
        b = r.config["input_beam"]
        assert 1e5 <= b["n_particle"] <= 1e7
        assert 20e-6 <= b["sigma_x"] <= 100e-6 and -20 <= b["chirp"] <= 0
        assert r.config["particle_deposition"]["xbins"] in (64, 128, 256)
        assert r.config["CSR_computation"]["xbins"] in (32, 64, 128)
        assert 100 <= r.steps <= 800
        fp = request_footprint(r.config, r.steps)
        assert fp > r.steps * 600 * 600 * 80


def test_admission_on_computed_footprints():
    """Requests admitted on the footprints request_footprint computes (not
    given ones): against a budget that holds the larger of two requests but
    not both, the second waits until the first retires."""
    reqs = mix_requests(n=2, rate=1000.0, seed=11, steps_range=(20, 30))
    fps = [request_footprint(r.config, r.steps) for r in reqs]
    log = []
    srv = _server(budget=max(fps) + min(fps) - 1, max_active=8, log=log)
    for r in reqs:
        assert srv.submit(Request(r.rid, r.config, 0.0, r.steps))
    assert [q.footprint for q in srv.waiting] == fps
    srv.tick(0.0)
    assert [s.req.rid for s in srv.active] == [reqs[0].rid] and srv.used == fps[0]
    rep = srv.run(simulated_clock=True)
    assert rep["finished"] == 2 and srv.used == 0


def test_regrid_past_the_admitted_grid_fails_the_request():
    """A history whose re-grid (int(500 sigma_max / sigma_min) per side,
    deposit.py:355-363) would pass the grid it was admitted for raises, and
    the server retires that request as failed, freeing its memory."""
    from pydfcsr_amd._lib import DfcsrError
    from pydfcsr_amd.deposit import DF_tracker, SliceMeta
    from pydfcsr_amd.serving import grid_bound
    tr = DF_tracker(dict(xbins=32, zbins=32), device="cpu")
    tr.max_grid = grid_bound({"particle_deposition": {}})
    assert tr.max_grid == 2000
    tr.max_grid = 600
    for t, sx in ((0.0, 10e-6), (0.1, 20e-6)):
        tr.cur = SliceMeta(t, sx, 50e-6, 0.0, 0.0, 32, 32, -5 * sx, 5 * sx, -2.5e-4, 2.5e-4,
                           10.0, 7)
        tr.append_DF()
    with pytest.raises(DfcsrError, match="admitted"):
        tr.append_interpolant(float("inf"), 4)

    class FailingSim(StubSim):
        def iter_phases(self, external_wake=True):
            yield ("wake", 1)
            yield ("done", 1)
            raise DfcsrError("re-grid to 1000 x 500 exceeds the 600 per side this history "
                             "was admitted for")

    srv = Server(device="cpu", hbm_budget=100, max_active=4, clock=Clock(),
                 sim_factory=lambda r: (FailingSim(r, []) if r.rid == 0 else StubSim(r, [])))
    srv.submit(Request(0, {}, 0.0, 3, footprint=60))
    srv.submit(Request(1, {}, 0.0, 2, footprint=60))
    rep = srv.run(simulated_clock=True)
    assert rep["failed"] == 1 and rep["finished"] == 1 and srv.used == 0
    assert "admitted" in srv.failed[0]
