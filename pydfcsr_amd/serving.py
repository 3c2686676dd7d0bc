"""Serving: continuous batching of many CSR simulations on one GPU.

Requests (a simulation config + arrival time + number of lattice steps) wait
in a FIFO queue.  At every tick the server

  1. admits arrived requests while a slot is free and the HBM footprint of the
     request (particles + history ring + scratch, an upper bound fixed at
     admission) fits the budget; each admitted simulation keeps its growing
     history on the device for its whole life;
  2. advances every active simulation through the host side of one step
     (tracking, deposition, density functions, history update, mesh) up to
     its wake;
  3. runs the wakes of all of them in ONE batched launch
     (dfcsr_wake_batched_launch: workgroups of (mesh point, phase) over all
     simulations), replayed from a HIP graph captured once per active-set
     shape;
  4. resumes each simulation to apply its kicks, retires finished ones and
     frees their memory, so waiting requests get in at the next tick.

With overlap (the default for real simulations) each active simulation runs
its host steps in a worker thread on a HIP stream of its own, so one
simulation's host work and device readbacks overlap the others' tracking,
deposition and re-grid kernels; the batched wake waits for every
simulation's stream by events and each simulation's kicks wait for the
wake.

Each simulation's wake is computed by the same per-workgroup code whether it
runs alone or in a mix, so its results do not depend on its neighbours.
Step latency is the wall time of the tick that completed the step; request
latency runs from a request's arrival to the end of its last step; the
report gives p50/p99 of both with throughput (simulation steps/s and
integrand points/s).
"""
from __future__ import annotations

import collections
import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Request:
    rid: int
    config: dict
    arrival: float
    steps: int
    precision: str = "mixed"
    footprint: int = 0


@dataclass
class _Slot:
    req: Request
    sim: object
    gen: object
    admitted_at: float
    steps_done: int = 0
    wake_results: list = field(default_factory=list)
    last_wake: tuple = None
    stream: object = None
    finished_at: float = None


# interp-grid side a request without an integer upper_limit is admitted for
# (the reference's chicane example caps it at 2000); a re-grid past it fails
# the request instead of overrunning its admitted memory
DEFAULT_GRID_BOUND = 2000


def grid_bound(config: dict) -> int:
    """Largest interp grid side the request's history may reach: its
    upper_limit when that is an int (deposit.py:361-363), else
    DEFAULT_GRID_BOUND."""
    ul = (config.get("particle_deposition") or {}).get("upper_limit")
    return int(ul) if isinstance(ul, int) else DEFAULT_GRID_BOUND


def request_footprint(config: dict, steps: int, precision: str = "mixed") -> int:
    """Upper bound of the device bytes one simulation holds (admission); the
    simulation's history is held to grid_bound(config) per side."""
    beam = config["input_beam"]
    n = int(beam.get("n_particle", 0)) or 1_000_000
    dep = config.get("particle_deposition") or {}
    nb = max(int(dep.get("xbins", 100)), int(dep.get("zbins", 100)), 100)
    grid = grid_bound(config)
    slots = steps + 2   # window and pending slice
    # per slot and grid node: the corner-packed record (160 B in float64, 80 B
    # in float32); per slot and cell of the cell-major rows (grid + 1 per x
    # row) the cell (40 B / 20 B, in float32 also its time difference, 20 B)
    rec, cell = (80, 2 * 4 * _lib.CELL) if precision == "mixed" else (160, 8 * _lib.CELL)
    hist = slots * grid * (grid * rec + (grid + 1) * cell)
    raw = slots * 5 * nb * nb * 8
    particles = n * 8 * (6 * 2 + 4)           # coordinates, tracking temporary, scratch
    deposit = 4 * nb * nb * 8 * 2 + 3 * nb * nb * 8
    comp = config.get("CSR_computation") or {}
    mesh = int(comp.get("xbins", 20)) * int(comp.get("zbins", 30)) * 8 * 4
    return int(hist + raw + particles + deposit + mesh + (64 << 20))


class BatchedWake:
    """One batched launch for the wakes of several simulations, optionally
    replayed from a captured HIP graph (one per launch shape).  The task
    table lives in a fixed device buffer rewritten before every launch; the
    partial sums in a device buffer that only grows (a reallocation drops the
    captured graphs, which hold its address)."""

    def __init__(self, device, precision="mixed", use_graph=True, max_tasks=256,
                 graph_cache=16):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.prec = _lib.DFCSR_PREC_MIXED if precision == "mixed" else _lib.DFCSR_PREC_FP64
        self.use_graph = use_graph
        self.lib = _lib.load()
        self.max_tasks = max_tasks
        nbytes = int(self.lib.dfcsr_wake_task_table_bytes(max_tasks))
        # two pinned host tables used in turn: a table is packed again only
        # after the copy that last read it has completed (its event)
        self.hosts = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self.copied = [None, None]
        self.turn = 0
        self.dev = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.part = torch.empty(0, dtype=torch.uint8, device=self.device)
        self.graphs = collections.OrderedDict()
        self.graph_cache = graph_cache
        self.seen = set()
        self.launches = 0
        self.replays = 0

    def run(self, items):
        """items: list of (History, Lattice, WakeParams, out{dE_dct, x_kick})."""
        torch = self.torch
        k = len(items)
        if k == 0:
            return
        if k > self.max_tasks:
            raise _lib.DfcsrError(f"{k} tasks > {self.max_tasks}")
        H = (_lib.History * k)(*[it[0] for it in items])
        L = (_lib.Lattice * k)(*[it[1] for it in items])
        P = (_lib.WakeParams * k)(*[it[2] for it in items])
        dE = (_lib.c_dp * k)(*[it[3]["dE_dct"].data_ptr() for it in items])
        xk = (_lib.c_dp * k)(*[it[3]["x_kick"].data_ptr() for it in items])
        batch = _lib.WakeBatch()
        i = self.turn
        self.turn ^= 1
        host = self.hosts[i]
        if self.copied[i] is not None:
            self.copied[i].synchronize()
        _lib.check(self.lib.dfcsr_wake_tasks_pack(k, H, L, P, self.prec, dE, xk,
                                                  host.data_ptr(), host.numel(),
                                                  ctypes.byref(batch)), "wake_tasks_pack")
        if batch.partial_bytes > self.part.numel():
            # freeing the old buffer is stream-ordered: torch's allocator hands it
            # out again only to work queued after the previous launch
            self.part = torch.empty(max(int(batch.partial_bytes), 2 * self.part.numel()),
                                    dtype=torch.uint8, device=self.device)
            self.graphs.clear()
            self.seen.clear()
        nbytes = int(self.lib.dfcsr_wake_task_table_bytes(k))
        # the previous launch may still read the device task buffer: the copy
        # is ordered after it on the same stream
        stream = torch.cuda.current_stream(self.device)
        self.dev[:nbytes].copy_(host[:nbytes], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        self.copied[i] = ev
        key = (k, batch.blocks_far, batch.blocks_near, batch.observers)

        def launch(s):
            _lib.check(self.lib.dfcsr_wake_batched_launch(self.dev.data_ptr(), ctypes.byref(batch),
                                                          self.prec, self.part.data_ptr(),
                                                          self.part.numel(), s),
                       "batched launch")

        if self.use_graph and key in self.seen:
            g = self.graphs.get(key)
            if g is None:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    launch(torch.cuda.current_stream(self.device).cuda_stream)
                self.graphs[key] = g
                if len(self.graphs) > self.graph_cache:
                    self.graphs.popitem(last=False)
            else:
                self.graphs.move_to_end(key)
            g.replay()
            self.replays += 1
        else:
            # first use of a shape runs eagerly (module load stays out of capture)
            launch(stream.cuda_stream)
            self.seen.add(key)
        self.launches += 1


class Server:
    """Continuous-batching request loop (module docstring)."""

    def __init__(self, device="cuda", hbm_budget=None, max_active=64, precision="mixed",
                 use_graph=True, sim_factory=None, clock=None, record_wakes=False,
                 overlap=None, workers=8):
        """overlap: each active simulation runs its host step in a worker
        thread on a stream of its own (default: on for real simulations), so
        one simulation's host work and readbacks overlap other simulations'
        kernels; the batched wake waits for all of them by events."""
        self.device = device
        self.precision = precision
        self.max_active = max_active
        self.record_wakes = record_wakes
        self.overlap = (sim_factory is None) if overlap is None else bool(overlap)
        self.pool = None
        if self.overlap:
            from concurrent.futures import ThreadPoolExecutor
            self.pool = ThreadPoolExecutor(max_workers=max(1, min(int(workers), 16)))
        self.waiting = collections.deque()
        self.active: list[_Slot] = []
        self.used = 0
        self.clock = clock or time.perf_counter
        self.sim_factory = sim_factory or self._make_sim
        if hbm_budget is None:
            import torch
            free, _total = torch.cuda.mem_get_info(torch.device(device))
            hbm_budget = int(0.85 * free)
        self.budget = int(hbm_budget)
        self.batcher = None if sim_factory is not None else BatchedWake(device, precision,
                                                                         use_graph)
        self.latencies = []          # per completed simulation step (s)
        self.queue_waits = []        # arrival -> admission (s)
        self.points = 0
        self.steps = 0
        self.ticks = 0
        self.finished = {}
        self.failed = {}
        self.rejected = []
        self.active_samples = []

    def _make_sim(self, req: Request):
        """A request's simulation.  Per-step wake files are written only when
        its CSR_computation asks for them explicitly (the reference's default
        is on; each write waits for the step's wake)."""
        from .csr import CSR2D
        cfg = dict(req.config)
        comp = dict(cfg.get("CSR_computation") or {})
        comp.setdefault("write_wakes", False)
        cfg["CSR_computation"] = comp
        sim = CSR2D(config=cfg, device=self.device, precision=self.precision,
                    keep_fp32=(self.precision == "mixed"), history_capacity=req.steps + 2)
        sim.DF_tracker.max_grid = grid_bound(cfg)
        return sim

    def submit(self, req: Request):
        if not req.footprint:
            req.footprint = request_footprint(req.config, req.steps, self.precision)
        if req.footprint > self.budget:
            self.rejected.append(req.rid)   # can never fit
            return False
        self.waiting.append(req)
        return True

    def _admit(self, now):
        while self.waiting and self.waiting[0].arrival <= now:
            req = self.waiting[0]
            if len(self.active) >= self.max_active or self.used + req.footprint > self.budget:
                break
            self.waiting.popleft()
            sim = self.sim_factory(req)
            gen = sim.iter_phases(external_wake=True)
            slot = _Slot(req, sim, gen, now)
            if self.overlap:
                import torch
                slot.stream = torch.cuda.Stream(torch.device(self.device))
                slot.stream.wait_stream(torch.cuda.current_stream(torch.device(self.device)))
            self.active.append(slot)
            self.used += req.footprint
            self.queue_waits.append(now - req.arrival)

    def _retire(self, slot: _Slot):
        self.active.remove(slot)
        self.used -= slot.req.footprint
        if slot.req.rid not in self.failed:
            self.finished[slot.req.rid] = slot
        if not self.record_wakes:
            slot.sim = None
            slot.gen = None

    def _on_stream(self, slot, fn):
        if slot.stream is None:
            return fn()
        import torch
        with torch.cuda.stream(slot.stream):
            return fn()

    def _advance(self, slot):
        """The host side of one step up to its wake (or the whole step when it
        has none): ("wake", (H, L, P, out), event), ("done",), ("end",) or
        ("fail", message)."""
        def step():
            try:
                kind, _step = next(slot.gen)
            except StopIteration:
                return ("end",)
            except _lib.DfcsrError as e:   # e.g. its history outgrew the admitted grid
                return ("fail", str(e))
            if kind != "wake":
                return ("done",)
            item = slot.sim.prepare_wake()
            ev = None
            if slot.stream is not None:
                import torch
                ev = torch.cuda.Event()
                ev.record(slot.stream)
            return ("wake", item, ev)
        return self._on_stream(slot, step)

    def _resume(self, slot, wake_done):
        """The rest of the step after its wake (kicks, statistics)."""
        def step():
            if wake_done is not None:
                slot.stream.wait_event(wake_done)
            try:
                next(slot.gen)
            except _lib.DfcsrError as e:
                return ("fail", str(e))
            if self.record_wakes:
                slot.wake_results.append((slot.sim.dE_dct.detach().clone(),
                                          slot.sim.x_kick.detach().clone()))
            return ("done",)
        return self._on_stream(slot, step)

    def _map(self, fn, items):
        if self.pool is None or len(items) < 2:
            return [fn(x) for x in items]
        return list(self.pool.map(fn, items))

    def tick(self, now):
        """One scheduling round; returns the number of steps completed."""
        self._admit(now)
        if not self.active:
            return 0
        t0 = self.clock()
        slots = list(self.active)
        need, done, ended = [], [], []
        for slot, r in zip(slots, self._map(self._advance, slots)):
            if r[0] == "end":
                ended.append(slot)
            elif r[0] == "fail":
                self.failed[slot.req.rid] = r[1]
                self._retire(slot)
            elif r[0] == "wake":
                need.append((slot, r[1], r[2]))
            else:
                done.append(slot)
        if need:
            items = [it for _, it, _ in need]
            if self.record_wakes:   # the inputs of each one's latest wake
                for (s, _, _), it in zip(need, items):
                    s.last_wake = (_lib.History.from_buffer_copy(it[0]),
                                   _lib.Lattice.from_buffer_copy(it[1]),
                                   _lib.WakeParams.from_buffer_copy(it[2]))
            wake_done = None
            if self.batcher is not None:
                torch = self.batcher.torch
                main = torch.cuda.current_stream(torch.device(self.device))
                for _, _, ev in need:
                    if ev is not None:
                        main.wait_event(ev)
                self.batcher.run(items)
                if self.overlap:
                    wake_done = torch.cuda.Event()
                    wake_done.record(main)
            else:
                for (s, _, _), it in zip(need, items):
                    s.sim.external_wake(it)
            resumed = [s for s, _, _ in need]
            for s, r in zip(resumed, self._map(lambda sl: self._resume(sl, wake_done), resumed)):
                self.points += s.sim.integrand_points()
                if r[0] == "fail":
                    self.failed[s.req.rid] = r[1]
                    self._retire(s)
                else:
                    done.append(s)
        if self.batcher is not None:
            self.batcher.torch.cuda.synchronize()
        t1 = self.clock()
        dt = t1 - t0
        for s in done:
            s.steps_done += 1
            self.latencies.append(dt)
            self.steps += 1
            if s.steps_done >= s.req.steps:
                ended.append(s)
        self.active_samples.append(len(self.active))
        for s in ended:
            if s in self.active:
                s.finished_at = now + dt   # in the run's time frame, as arrivals
                self._retire(s)
        self.ticks += 1
        return len(done)

    def run(self, duration=None, max_ticks=None, simulated_clock=False, tick_dt=0.0):
        """Runs until every request finished (or duration / max_ticks).
        simulated_clock: arrivals are compared with a virtual time advanced
        by each tick's wall time (plus tick_dt), not with real time."""
        start = self.clock()
        vnow = 0.0
        while self.waiting or self.active:
            now = vnow if simulated_clock else self.clock() - start
            if duration is not None and (self.clock() - start) > duration:
                break
            if max_ticks is not None and self.ticks >= max_ticks:
                break
            if not self.active and self.waiting and self.waiting[0].arrival > now:
                if simulated_clock:
                    vnow = self.waiting[0].arrival
                else:
                    time.sleep(min(0.05, self.waiting[0].arrival - now))
                continue
            t0 = self.clock()
            self.tick(now)
            vnow += (self.clock() - t0) + tick_dt
        elapsed = self.clock() - start
        return self.report(elapsed)

    def report(self, elapsed):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        qw = np.array(self.queue_waits) if self.queue_waits else np.zeros(1)
        done = [sl for sl in self.finished.values() if sl.finished_at is not None]
        rl = np.array([sl.finished_at - sl.req.arrival for sl in done]) if done else np.zeros(1)
        return {
            "elapsed_s": round(elapsed, 3),
            "ticks": self.ticks,
            "sim_steps": self.steps,
            "steps_per_s": round(self.steps / elapsed, 3) if elapsed > 0 else None,
            "gpts_per_s": round(self.points / elapsed / 1e9, 3) if elapsed > 0 else None,
            "step_latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "step_latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "queue_wait_s_p50": round(float(np.percentile(qw, 50)), 3),
            "queue_wait_s_p99": round(float(np.percentile(qw, 99)), 3),
            "request_latency_s_p50": round(float(np.percentile(rl, 50)), 3),
            "request_latency_s_p99": round(float(np.percentile(rl, 99)), 3),
            "finished": len(self.finished),
            "failed": len(self.failed),
            "rejected": len(self.rejected),
            "still_waiting": len(self.waiting),
            "still_active": len(self.active),
            "mean_active": round(float(np.mean(self.active_samples)), 2)
            if self.active_samples else 0.0,
            "batched_launches": self.batcher.launches if self.batcher else 0,
            "graph_replays": self.batcher.replays if self.batcher else 0,
            "hbm_budget_bytes": self.budget,
        }


# -- the request mix ------------------------------------------------------------------------
def make_request(rid, arrival, steps, n_particle, grid, mesh, lattice, strength, sigma_x,
                 sigma_z, chirp, seed, upper_limit=600, Nxz=200, precision="mixed"):
    """One request: a Gaussian beam (1 nC, 1 GeV, energy spread 1e-4) through
    a single 0.5 m dipole of radius `strength` [m] (lattice 'dipole') or a
    4-bend chicane (0.3 m bends at +, -, -, + `strength` rad, drifts 2/1/2 m;
    lattice 'chicane'), in `steps` steps; deposition grid `grid` per side
    (SG 7/2), wake mesh `mesh` per side, integration Nxz x Nxz, history grid
    capped at upper_limit per side."""
    from collections import OrderedDict
    lat = OrderedDict()
    if lattice == "dipole":
        L = 0.5
        lat["step_size"] = L / steps
        lat["B1"] = dict(type="dipole", L=L, angle=L / strength, nsep=1)
    else:
        ang = strength
        total = 4 * 0.3 + 5.0
        lat["step_size"] = total / steps
        lat["B1"] = dict(type="dipole", L=0.3, angle=ang, E1=0.0, E2=ang, nsep=1)
        lat["D1"] = dict(type="drift", L=2.0, nsep=1)
        lat["B2"] = dict(type="dipole", L=0.3, angle=-ang, E1=-ang, E2=0.0, nsep=1)
        lat["D2"] = dict(type="drift", L=1.0, nsep=1)
        lat["B3"] = dict(type="dipole", L=0.3, angle=-ang, E1=0.0, E2=-ang, nsep=1)
        lat["D3"] = dict(type="drift", L=2.0, nsep=1)
        lat["B4"] = dict(type="dipole", L=0.3, angle=ang, E1=ang, E2=0.0, nsep=1)
    cfg = {
        "input_beam": dict(style="gaussian", n_particle=int(n_particle), charge=1e-9,
                           energy=1e9, sigma_x=sigma_x, sigma_z=sigma_z, chirp=chirp,
                           sigma_delta=1e-4, seed=int(seed)),
        "input_lattice": dict(lattice_input_file=f"<request {rid}>", _config=lat),
        "particle_deposition": dict(xbins=int(grid), zbins=int(grid), xlim=5, zlim=5,
                                    filter_order=2, filter_window=7, velocity_threhold=1000,
                                    upper_limit=upper_limit),
        "CSR_integration": dict(n_formation_length=4, zbins=Nxz, xbins=Nxz),
        "CSR_computation": dict(xbins=int(mesh), zbins=int(mesh), xlim=5, zlim=5, apply_CSR=1,
                                transverse_on=1, write_wakes=False),
    }
    return Request(rid, cfg, arrival, int(steps), precision)


def mix_requests(n=64, rate=2.0, seed=0, steps_range=(100, 800), particles_range=(1e5, 1e7),
                 grids=(64, 128, 256), meshes=(32, 64, 128), upper_limit=600, Nxz=200,
                 precision="mixed"):
    """Seeded request stream: Poisson arrivals at `rate` per second; particle
    count log-uniform; deposition grid, wake mesh, lattice (single dipole with
    rho 5-20 m, or a 4-bend chicane at 0.02-0.08 rad), history length, beam
    sizes (20-100 um) and chirp (-20..0 /m) drawn per request (make_request)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        npart = int(10 ** rng.uniform(math.log10(particles_range[0]),
                                      math.log10(particles_range[1])))
        nb = int(rng.choice(grids))
        mesh = int(rng.choice(meshes))
        if rng.random() < 0.5:
            kind, strength = "dipole", rng.uniform(5.0, 20.0)
        else:
            kind, strength = "chicane", rng.uniform(0.02, 0.08)
        sx = rng.uniform(20e-6, 100e-6)
        sz = rng.uniform(20e-6, 100e-6)
        chirp = rng.uniform(-20.0, 0.0)
        bseed = int(rng.integers(1 << 30))
        out.append(make_request(rid, t, steps, npart, nb, mesh, kind, strength, sx, sz, chirp,
                                bseed, upper_limit, Nxz, precision))
    return out
