"""Continuous-batching request loop: k-eigenvalue solves of branch cases
served on one device through the batch engine.

Requests arrive over time. A request is admitted as soon as the batch's
device-memory budget has room for its state (FIFO); every admitted problem
keeps its flux, sources and boundary fluxes on the device until it finishes.
The device advances all admitted problems together a few outer iterations
at a time; finished problems leave the batch, and waiting ones join at the
next iteration boundary. A finished problem's handle goes back to the
factory, which gives it the next case of the same geometry and group count
(mocgpu_reload_xs: new cross sections, no device allocation or free); the
budget counts the problems in the batch.

While the device works on one chunk, the host does only host work for the
requests that have arrived (their cross-section libraries). Their device
handles are made once the chunk has finished: with more streams than the
device's hardware queues (GPU_MAX_HW_QUEUES, 4 by default), a handle's stream
can share the batch stream's queue, and its set-up copies would wait behind
the running chunk.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .branch import BranchFactory, Case
from .sweeper import CapacityError, GpuBatchSet


@dataclass
class Record:
    index: int
    case: Case
    t_arrive: float
    t_admit: float = float("nan")
    t_done: float = float("nan")
    k: float = float("nan")
    iterations: int = 0
    converged: bool = False
    flux: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (seconds from 0) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


class Server:
    def __init__(self, factory: BranchFactory, budget_bytes: int = 0, k_tol: float = 1e-7,
                 psi_tol: float = 1e-6, max_iter: int = 3000, chunk: int = 1, keep_flux: bool = False,
                 lanes: int = 1):
        """chunk: outer iterations per graph replay between admissions. One
        admits arrivals and releases finished problems soonest. On requests
        from pin, 3x3, 17x17 and 2x2-of-17x17 lattices arriving at half the
        burst rate, it gave p50 464 ms against 688 ms with 4, at a slightly
        higher throughput. lanes: batches (each on its own stream) the
        requests are spread over. Two raised the 2000-request burst from 109
        to 115 solves/s but doubled the latency at 0.5x load (p50 483 ->
        905 ms): a coarse solve shares its CU with the other batch's sweeps
        and the slowest one sets every request's pace, so one is the
        default."""
        self.factory = factory
        self.batch = GpuBatchSet(lanes, budget_bytes=budget_bytes, use_graph=True)
        self.k_tol, self.psi_tol, self.max_iter = k_tol, psi_tol, max_iter
        self.chunk = chunk
        self.keep_flux = keep_flux

    def serve(self, arrivals, cases: list[Case]) -> list[Record]:
        recs = [Record(i, c, float(t)) for i, (t, c) in enumerate(zip(arrivals, cases))]
        future = deque(range(len(recs)))  # not yet arrived
        staged = deque()                  # arrived: (index, cross-section library)
        ready = deque()                   # (index, device handle) waiting for room
        active = {}                       # batch id -> (index, handle)
        b = self.batch
        t0 = time.perf_counter()
        busy = False
        while future or staged or ready or active:
            now = time.perf_counter() - t0
            # host work while the device runs the previous chunk
            while future and recs[future[0]].t_arrive <= now:
                i = future.popleft()
                staged.append((i, self.factory.stage(recs[i].case)))
            if busy:
                for cid, st in b.poll():
                    i, s = active.pop(cid)
                    r = recs[i]
                    r.t_done = time.perf_counter() - t0
                    r.k, r.iterations, r.converged = st.k, st.iterations, bool(st.converged)
                    if self.keep_flux:
                        r.flux = s.flux()
                    b.remove(cid)
                    self.factory.release(r.case, s)
                busy = False
            # the device is idle: make the handles of the staged requests
            while staged:
                i, xsl = staged.popleft()
                ready.append((i, self.factory.make(recs[i].case, xsl)))
            while ready:
                i, s = ready[0]
                try:
                    cid = b.add(s, self.k_tol, self.psi_tol, self.max_iter)
                except CapacityError:
                    if not active:
                        raise  # larger than the whole budget: it can never be admitted
                    break
                ready.popleft()
                active[cid] = (i, s)
                recs[i].t_admit = time.perf_counter() - t0
            if active:
                b.run(self.chunk)
                busy = True
            elif future:
                time.sleep(max(0.0, recs[future[0]].t_arrive - (time.perf_counter() - t0)))
        return recs

    def close(self):
        self.batch.close()


def serve_serial(factory: BranchFactory, arrivals, cases: list[Case], k_tol=1e-7, psi_tol=1e-6,
                 max_iter=3000) -> list[Record]:
    """The plain alternative: requests solved one at a time in arrival order
    with the single-problem device path (mocgpu_eigen_solve)."""
    recs = [Record(i, c, float(t)) for i, (t, c) in enumerate(zip(arrivals, cases))]
    t0 = time.perf_counter()
    for r in recs:
        now = time.perf_counter() - t0
        if r.t_arrive > now:
            time.sleep(r.t_arrive - now)
        r.t_admit = time.perf_counter() - t0
        s = factory.make(r.case)
        rc, st = s.eigen_solve(k_tol, psi_tol, max_iter)
        r.t_done = time.perf_counter() - t0
        r.k, r.iterations, r.converged = st.k, st.iterations, bool(st.converged)
        factory.release(r.case, s)
    return recs


def summarize(recs: list[Record]) -> dict:
    lat = np.array([r.latency for r in recs])
    span = max(r.t_done for r in recs) - min(r.t_arrive for r in recs)
    return {
        "requests": len(recs),
        "solves_per_s": len(recs) / span if span > 0 else float("nan"),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "converged": int(sum(r.converged for r in recs)),
    }
