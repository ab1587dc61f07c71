#!/usr/bin/env python3
"""Benchmark: MOC segment-group updates/s of the gfx950 sweep on one MI355X.

Workload: the 17x17 assembly (25 moderator guide tubes), 7 groups, Chebyshev-
Yamamoto 32 azimuths x 3 polar per octant, ray spacing 0.03 cm, reflective
(seeded synthetic cross sections, mocc_amd.lattice_gen.c2_deck).

A step is one inner sweep of all 7 groups (sweep kernel + flux reduction),
i.e. 2 * P * S * G segment-group updates. Inputs are resident in HBM before
the timed region. The JSON line also carries the roofline of the sweep
kernel (HIP-event kernel time on the library's stream) and the FP64 CPU
oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MOC segment-group updates/s; batched lattice solves/s + p99 at keff err≤1e-5"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9  # CUs x SIMDs x lanes/clk x clock (non-packed FP32)


def algorithmic_bytes(p) -> int:
    """Unique data one all-group inner sweep must touch: segment stream
    (int32 FSR id + fp32 length), Sigma_t and q read plus phi written (FP32
    words, 16 B per FSR-group) and the boundary angular flux read and written
    (fp32) for both directions of every ray, polar angle and group."""
    S, F, G, R, P = p.n_seg, p.n_fsr, p.n_group, p.n_ray, p.n_polar
    return 8 * S + 16 * F * G + 16 * R * P * G


def read_traffic(path: str):
    """HBM bytes per sweep-kernel dispatch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_dispatch"])
    except (KeyError, ValueError):
        return None


def cpu_baseline(hp, n_sweeps: int):
    """FP64 oracle (reference loop structure, OpenMP) timed on this host."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    o = Oracle(hp, n_inner=1, gs_boundary=True, energy_mode=0, n_threads=threads)
    o.initialize()
    o.calc_fission_source(1.0)
    for g in range(hp.n_group):
        o.build_source(g)
    ms = o.time_sweeps(n_sweeps)
    ups = hp.updates_per_sweep() / (ms * 1e-3)
    return {"value": ups, "unit": "updates/s", "cores": threads, "kind": "port",
            "sample": f"median of {n_sweeps} all-group inner sweeps of the same 17x17/7-group "
                      f"deck (after 1 warm-up), FP64, sweep1g only, OMP threads={threads}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-sweeps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs on one GPU only "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    from mocc_amd.sweeper import GpuMoCSweeper

    deck_dir = tempfile.mkdtemp(prefix="mocc_bench_")
    path = lg.write_config("c2", deck_dir)
    hp = HostProblem.load(path)
    p = hp.p
    dev = GpuMoCSweeper(hp, n_inner=1, energy_mode=1)
    dev.initialize()
    dev.calc_fission_source(1.0)
    dev.store_old_flux()
    for g in range(hp.n_group):
        dev.build_source(g)
    U = dev.updates_per_sweep()

    # warm-up and timed steps: one all-group inner sweep each
    from mocc_amd.native import gpu
    L = gpu()
    for _ in range(args.warmup):
        dev.sweep_all_groups()
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.sweep_all_groups()
    dev.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = U / (ms_per_step * 1e-3)

    # kernel-only time with HIP events on the library stream
    _, kernel_ms = dev.time_sweeps(max(10, min(args.steps, 50)))
    B = algorithmic_bytes(p)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", "sweep_c2_pmc.txt"))
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_sweep<3,false>", "kernel_ms": kernel_ms, "algorithmic_bytes": B,
                "valu_frac_est": (U / (kernel_ms * 1e-3)) * 10.0 / VALU_PEAK_LANE_OPS}
    line = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "17x17 assembly (25 guide tubes), 7 groups, CY 32x3, spacing 0.03 cm, "
                               "reflective; one all-group inner sweep per step",
                   "n_fsr": p.n_fsr, "n_seg": p.n_seg, "n_ray": p.n_ray, "updates_per_step": U,
                   "parallelism": "single GPU"},
        "roofline": roofline,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(hp, args.cpu_sweeps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
