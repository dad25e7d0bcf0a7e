"""Serving: a request loop that admits solidification runs by free HBM and
advances every admitted run together (continuous batching over pf_batch_*).

A request is a Problem (grid, cooling rate, seeds, step count, precision)
with an arrival time.  The server loop, one tick at a time:
  1. admits arrived requests in arrival order while the batch's HBM budget
     has room for their fields and the cells in flight stay under
     `max_active_cells`; a request that does not fit waits, and a later one
     that fits may overtake it, at most `overtake_limit` times in a row (then
     the head blocks admission until it fits).  The cap matters for latency:
     every admitted run advances by the same number of steps per tick, so
     beyond the cell count that saturates the GPU more runs in flight only
     stretch each of them;
  2. enqueues up to `tick_steps` steps of every admitted run (one batched
     launch per sweep for all of them; runs that finish mid-tick stop at their
     own step count on the device) -- the tick ends at the first completion
     when `end_at_completion`, else it runs its full length -- and, while the
     GPU works, builds the initial fields of requests that have arrived (host
     work that would otherwise leave the GPU idle at admission);
  3. retires the runs that completed: downloads phi/c/T and the diagnostics,
     frees their memory, and records the completion time.
Requests that arrive while others are on the GPU are therefore admitted as
memory frees up and their work overlaps with the runs already in flight.
The sequential baseline serves the same stream one request at a time
through the single-instance loop (pf_run_steps).

Time is the host's wall clock; arrivals are replayed in real time (the loop
sleeps when it has nothing to run and the next arrival is in the future).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import capi
from .ensemble import Problem, draw_problem


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    problem: Problem


@dataclass
class Completion:
    rid: int
    arrival: float
    admitted: float
    done: float
    fields: Optional[tuple] = None  # (phi, c, T) if kept
    diag: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class ServeReport:
    completions: List[Completion]
    makespan: float
    busy_ticks: int = 0
    peak_bytes: int = 0
    peak_instances: int = 0
    gpu_wait_s: float = 0.0  # host time spent waiting for ticks to finish
    overtakes: int = 0
    # host time in admissions (pf_batch_add and tuning), in enqueueing ticks
    # (pf_batch_run, which rebuilds the tables and recaptures the graph after
    # a change of the instance set) and in retiring (download and remove)
    admit_s: float = 0.0
    enqueue_s: float = 0.0
    retire_s: float = 0.0
    prepare_s: float = 0.0  # building initial fields while a tick runs
    # one (start, end, steps, runs in flight, cells in flight, enqueue time)
    # per tick
    tick_log: List[tuple] = field(default_factory=list)

    def latencies(self):
        return np.array([c.latency for c in self.completions])

    def summary(self) -> Dict[str, float]:
        lat = self.latencies()
        return {"requests": len(self.completions),
                "solves_per_s": len(self.completions) / self.makespan,
                "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()),
                "max_s": float(lat.max()),
                "makespan_s": self.makespan,
                "peak_instances": self.peak_instances,
                "peak_bytes": self.peak_bytes,
                "ticks": self.busy_ticks,
                "gpu_wait_s": self.gpu_wait_s,
                "admit_s": self.admit_s,
                "enqueue_s": self.enqueue_s,
                "retire_s": self.retire_s,
                "prepare_s": self.prepare_s,
                "overtakes": self.overtakes}


def request_stream(n: int = 32, rate: float = 2.0, seed: int = 11, steps_lo: int = 1000,
                   steps_hi: int = 20000, edge_lo: int = 32, edge_hi: int = 256,
                   max_seeds: int = 16) -> List[Request]:
    """Poisson arrivals at `rate` per second; grid edges log-uniform in
    [edge_lo, edge_hi] per axis (z >= x, y), cooling 5-50 K/s, 1-16 seeds,
    step count uniform in [steps_lo, steps_hi], FP64 or FP32-mixed 50/50."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for q in range(n):
        t += float(rng.exponential(1.0 / rate))
        prec = capi.PF_FP64 if rng.random() < 0.5 else capi.PF_FP32_MIXED
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        out.append(Request(q, t, draw_problem(rng, edge_lo, edge_hi, max_seeds, prec, steps,
                                              tag=f"r{q}")))
    return out


def instance_bytes(problem: Problem) -> int:
    return 6 * problem.cells * (8 if problem.precision == capi.PF_FP64 else 4)


def serve_batched(requests: List[Request], hbm_budget: int = 0, tick_steps: int = 200,
                  graph_steps: int = 16, keep: bool = False, autotune: bool = True,
                  max_active_cells: Optional[int] = None, end_at_completion: bool = True,
                  overtake_limit: int = 8, lib=None,
                  log: Optional[Callable[[str], None]] = None) -> ServeReport:
    lib = lib or capi.load()
    b = capi.Batch(hbm_budget, lib=lib)
    try:
        b.set_graph_steps(graph_steps)
        _, budget = b.bytes()
        queue = sorted(requests, key=lambda r: r.arrival)
        active: Dict[int, tuple] = {}  # slot -> (request, admitted time)
        prepared: Dict[int, tuple] = {}  # rid -> (setup, initial fields) built ahead
        tuned = set()
        done: List[Completion] = []
        rep = ServeReport(done, 0.0)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        head_skips = 0
        while queue or active:
            # 1. admission by free HBM in arrival order, overtaking bounded
            q = 0
            while q < len(queue) and queue[q].arrival <= now():
                r = queue[q]
                used, _ = b.bytes()
                fits = used + instance_bytes(r.problem) <= budget
                in_flight = sum(x.problem.cells for x, _ in active.values())
                if fits and active and max_active_cells and \
                        in_flight + r.problem.cells > max_active_cells:
                    fits = False
                if not fits:
                    if not active and used + instance_bytes(r.problem) > budget:
                        raise RuntimeError(f"request {r.rid} needs more than the HBM budget")
                    if q == 0 and head_skips >= overtake_limit:
                        break  # the head has been overtaken enough: wait for it
                    q += 1
                    continue
                if q > 0 and head_skips >= overtake_limit:
                    break  # (also within one admission pass)
                ta = time.perf_counter()
                s, f = prepared.pop(r.rid, None) or r.problem.setup(lib)
                slot = b.add(s.p, s.L, r.problem.steps, f[0], f[2], f[4])
                del f
                active[slot] = (r, now())
                queue.pop(q)
                if q > 0:
                    head_skips += 1
                    rep.overtakes += 1
                else:
                    head_skips = 0
                if autotune:
                    cls = b.size_class(r.problem.precision, r.problem.deck.Nx, r.problem.deck.Ny)
                    if cls not in tuned:  # first instance of a size class: tune it
                        b.autotune(reps=1, size_class=cls)
                        tuned.add(cls)
                rep.admit_s += time.perf_counter() - ta
            used, _ = b.bytes()
            rep.peak_bytes = max(rep.peak_bytes, used)
            rep.peak_instances = max(rep.peak_instances, len(active))
            if not active:
                wait = queue[0].arrival - now()
                if wait > 0:
                    time.sleep(wait)
                continue
            # 2. advance every admitted run (asynchronously) and build the
            # inputs of arrived requests while the GPU works
            left = [b.status(sl)[1] for sl in active]
            k = min(tick_steps, min(left) if end_at_completion else max(left))
            ts = now()
            b.run(k)
            te = now() - ts
            rep.enqueue_s += te
            tp = time.perf_counter()
            for r in queue:
                if r.arrival > now():
                    break
                if r.rid not in prepared:
                    prepared[r.rid] = r.problem.setup(lib)
            rep.prepare_s += time.perf_counter() - tp
            tw = time.perf_counter()
            b.synchronize()
            rep.gpu_wait_s += time.perf_counter() - tw
            rep.busy_ticks += 1
            rep.tick_log.append((ts, now(), k, len(active),
                                 sum(x.problem.cells for x, _ in active.values()), te))
            # 3. retire the completed runs
            tr = time.perf_counter()
            for sl in [sl for sl in active if b.status(sl)[1] == 0]:
                r, adm = active.pop(sl)
                c = Completion(r.rid, r.arrival, adm, now())
                if keep:
                    c.fields = b.download(sl, r.problem.cells)
                    c.diag = b.diagnosticos(sl)
                b.remove(sl)
                done.append(c)
                if log:
                    log(f"[serve] request {r.rid} done: latency {c.latency:.2f} s, "
                        f"{len(active)} running, {len(queue)} waiting")
            rep.retire_s += time.perf_counter() - tr
        rep.makespan = now()
        return rep
    finally:
        b.close()


def serve_sequential(requests: List[Request], keep: bool = False, lib=None,
                     log: Optional[Callable[[str], None]] = None) -> ServeReport:
    """The same stream served one request at a time (FIFO) through the
    single-instance fused loop."""
    from .run import Simulation

    lib = lib or capi.load()
    done: List[Completion] = []
    rep = ServeReport(done, 0.0)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - now()
        if wait > 0:
            time.sleep(wait)
        adm = now()
        _, f = r.problem.setup(lib)
        pr = r.problem
        sim = Simulation(pr.deck, precision=pr.precision, seed=pr.noise_seed, fields=f, lib=lib)
        sim.run(pr.steps)
        sim.ctx.synchronize()
        c = Completion(r.rid, r.arrival, adm, now())
        if keep:
            g = sim.fields()
            c.fields = (g["phi_new"], g["c_new"], g["T_new"])
            c.diag = sim.diagnostics()
        sim.close()
        done.append(c)
        rep.peak_instances = 1
        if log:
            log(f"[serve-seq] request {r.rid} done: latency {c.latency:.2f} s")
    rep.makespan = now()
    return rep
