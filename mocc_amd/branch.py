"""Cross-section branch cases of a few lattice geometries.

A branch case is one lattice (a single pin, 3x3 pins, a 17x17 assembly or
2x2 assemblies) with 2 or 7 groups and every cross section of the seeded base
materials scaled by U[0.9, 1.1]. Geometries are traced once per (lattice,
ray spacing) and their device ray plan is uploaded once; each case then only
adds its own cross sections and state (mocgpu_create_like), so admitting a
case costs a library reload and a few small device copies.
"""
from __future__ import annotations

import os
import threading
from dataclasses import dataclass

import numpy as np

from . import lattice_gen as lg
from .native import HostProblem
from .sweeper import GpuMoCSweeper

N_AZIMUTHAL = 16  # per quadrant, CY with 3 polar angles
LATTICES = ("pin", "3x3", "17x17", "2x2x17x17")


def lattice_deck(kind: str, groups: int, spacing: float) -> lg.Deck:
    name = f"br_{kind}_{groups}g_{int(round(spacing * 1000))}"
    if kind == "pin":
        return lg.pin_lattice(1, groups, N_AZIMUTHAL, spacing, name)
    if kind == "3x3":
        return lg.pin_lattice(3, groups, N_AZIMUTHAL, spacing, name)
    if kind == "17x17":
        return lg.pin_lattice(17, groups, N_AZIMUTHAL, spacing, name, guide_tubes=True)
    if kind == "2x2x17x17":
        d = lg.Deck(name, groups, N_AZIMUTHAL, spacing)
        d.lattices = {1: (17, 17, lg.lattice_rows(17, 0, True))}
        d.core = [[1, 1], [1, 1]]
        return d
    raise ValueError(kind)


@dataclass
class Case:
    kind: str
    groups: int
    spacing: float
    seed: int  # branch seed (cross-section draws)


class BranchFactory:
    """Makes device handles for branch cases. Rays are traced once per
    (lattice, spacing); a prototype handle per geometry holds the device ray
    plan that every case of that geometry shares."""

    BASE_SEED = 4

    def __init__(self, work_dir: str, energy_mode: int = 1, n_inner: int = 1, use_cmfd: bool = False):
        """use_cmfd: CMFD-accelerated solves (mocc's CMFD decks use 5 inner
        iterations per outer; with 1 the coarse system can lose its
        diagonal dominance). Every geometry's ray plan uses the factory's band
        width and lane map (the device defaults until tune() picks them)."""
        self.work_dir = work_dir
        self.energy_mode = energy_mode
        self.n_inner = n_inner
        self.use_cmfd = use_cmfd
        self.band_rays = 0     # 0: the device default (64)
        self.lane_map = None   # None: the device default
        self.tuning: list = []  # (band rays, lane map, seconds) of the last tune()
        self._pool: dict = {}   # (lattice, spacing, groups) -> released handles
        self._hp: dict = {}
        self._proto: dict = {}
        self._lock = threading.Lock()
        os.makedirs(work_dir, exist_ok=True)

    def _host(self, kind: str, spacing: float) -> HostProblem:
        key = (kind, spacing)
        hp = self._hp.get(key)
        if hp is None:
            path = lg.write_deck(lattice_deck(kind, 2, spacing), self.work_dir, self.BASE_SEED)
            hp = HostProblem.load(path)
            self._hp[key] = hp
        return hp

    def deck(self, c: Case) -> lg.Deck:
        return lattice_deck(c.kind, c.groups, c.spacing)

    def stage(self, c: Case) -> str:
        """Host-only preparation of a case: trace its geometry if new and
        write its cross-section library; returns the library's path."""
        self._host(c.kind, c.spacing)
        d = self.deck(c)
        xsl = os.path.join(self.work_dir, f"{d.name}_b{c.seed}.xsl")
        lg.write_branch_xsl(d, self.BASE_SEED, c.seed, 0.1, xsl)
        return xsl

    def host_problem(self, c: Case, staged: str | None = None) -> HostProblem:
        """The traced geometry of the case, loaded with the case's cross
        sections (valid until the next call for the same geometry). staged:
        the library written by stage(c), consumed here."""
        xsl = staged if staged is not None else self.stage(c)
        hp = self._host(c.kind, c.spacing)
        hp.reload_xs(xsl)
        os.remove(xsl)
        return hp

    def make(self, c: Case, staged: str | None = None) -> GpuMoCSweeper:
        """Device handle of a case (device allocations and copies on the
        handle's own stream), or a released handle of the same geometry and
        group count given the case's cross sections (no allocation)."""
        with self._lock:
            hp = self.host_problem(c, staged)
            pool = self._pool.get((c.kind, c.spacing, c.groups))
            if pool:
                s = pool.pop()
                s.reload_xs(hp)
                return s
            gkey = (c.kind, c.spacing)
            proto = self._proto.get(gkey)
            if proto is None:
                proto = GpuMoCSweeper(hp, n_inner=self.n_inner, energy_mode=self.energy_mode,
                                      use_cmfd=self.use_cmfd, band_rays=self.band_rays, lane_map=self.lane_map)
                self._proto[gkey] = proto
            return GpuMoCSweeper(hp, n_inner=self.n_inner, energy_mode=self.energy_mode, use_cmfd=self.use_cmfd,
                                 like=proto)

    def tune(self, sample: list[Case], widths=(64, 48, 32), maps=(1, 0, 2), n_outer: int = 3) -> tuple[int, int]:
        """Band width x lane map for the batched workload, by timing it: for
        every candidate, all geometries get that layout, and n_outer outer
        iterations of one graph-replayed batch of the sample cases are timed
        on the device (after one untimed iteration that builds the graph).
        The fastest candidate is kept for the handles made afterwards.

        The layout is chosen for the batch as a whole, not per geometry: the
        batch runs every problem's bands in workgroups as wide as its widest
        plan, so a geometry tuned alone to 32-ray bands would run them with
        half its lanes idle beside 64-ray plans. Returns (band rays, lane map)."""
        import time

        from .sweeper import GpuBatch
        results = []
        for w in widths:
            for m in maps:
                self.close()
                self.band_rays, self.lane_map = w, m
                hs = [self.make(c) for c in sample]
                b = GpuBatch(use_graph=True)
                for h in hs:  # tolerances no problem meets: all run every iteration
                    b.add(h, 0.0, 0.0, n_outer + 2)
                b.run(1)
                b.poll()
                t0 = time.perf_counter()
                b.run(n_outer)
                b.poll()
                results.append((w, m, time.perf_counter() - t0))
                b.close()
                for h in hs:
                    h.close()
        self.close()
        self.tuning = results
        w, m, _ = min(results, key=lambda r: r[2])
        self.band_rays, self.lane_map = w, m
        return w, m

    def release(self, c: Case, s: GpuMoCSweeper):
        """Hand back the handle of a finished case for reuse by make()."""
        with self._lock:
            self._pool.setdefault((c.kind, c.spacing, c.groups), []).append(s)

    def tuned(self) -> dict:
        """(lattice, spacing) -> (band rays, lane map) of each geometry's plan."""
        out = {}
        for k, p in self._proto.items():
            o = p.options()
            out[k] = (int(o.band_rays), int(o.lane_map))
        return out

    def close(self):
        for pool in self._pool.values():
            for s in pool:
                s.close()
        self._pool.clear()
        for p in self._proto.values():
            p.close()
        self._proto.clear()


def c4_cases(n: int = 256, seed: int = 4) -> list[Case]:
    """The batched branch workload: a quarter single pins, half 3x3 lattices,
    a quarter 17x17 assemblies; 2 and 7 groups alternating; ray spacing drawn
    from {0.02, 0.035, 0.05} cm; per-case cross sections +-10 %."""
    rng = np.random.default_rng(seed)
    kinds = ["pin"] * (n // 4) + ["3x3"] * (n // 2) + ["17x17"] * (n - n // 4 - n // 2)
    out = []
    for i, k in enumerate(kinds):
        sp = float(rng.choice([0.02, 0.035, 0.05]))
        out.append(Case(k, 2 if i % 2 == 0 else 7, sp, 1000 + i))
    return out


def c5_request(rng: np.random.Generator, i: int) -> Case:
    """One serving request: a lattice of {pin, 3x3, 17x17, 2x2 of 17x17} with
    2 or 7 groups at 0.05 cm spacing and its own cross-section draws."""
    kind = LATTICES[int(rng.integers(0, len(LATTICES)))]
    return Case(kind, int(rng.choice([2, 7])), 0.05, 50000 + i)
