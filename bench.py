#!/usr/bin/env python3
"""Benchmark: MOC segment-group updates/s of the gfx950 sweep on one MI355X.

Workload: the 17x17 assembly (25 moderator guide tubes), 7 groups, Chebyshev-
Yamamoto 32 azimuths x 3 polar per octant, ray spacing 0.03 cm, reflective
(seeded synthetic cross sections, mocc_amd.lattice_gen.assembly17_deck).

A step is one inner sweep of all 7 groups (sweep kernel + flux reduction),
i.e. 2 * P * S * G segment-group updates. Inputs are resident in HBM before
the timed region. The JSON line also carries the roofline of the sweep
kernel (HIP-event kernel time on the library's stream), the FP64 CPU oracle
timed on this host (all threads and one), the batched branch-case solves/s
against one-at-a-time device solves ("batched", autotuned band layouts,
keff of a sample of every lattice kind against FP64), the serving loop's
p50/p99 latency under Poisson arrivals ("serving", 2000 requests per rate,
served keffs against FP64) and the full eigen solve of the 51x51-pin core
("core51", keff against the FP64 oracle's solve of the same deck).
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
# VALU issue: 256 CUs x 4 SIMDs, a wave64 VALU instruction occupies its SIMD
# for 4 cycles (16 lanes per cycle; longer for transcendentals), 2.4 GHz.
N_SIMD, VALU_CYCLES_PER_INST, CLOCK_HZ = 256 * 4, 4, 2.4e9


def algorithmic_bytes(p) -> int:
    """Unique data one all-group inner sweep must touch: segment stream
    (int32 FSR id + fp32 length), Sigma_t and q read plus phi written (FP32
    words, 16 B per FSR-group) and the boundary angular flux read and written
    (fp32) for both directions of every ray, polar angle and group."""
    S, F, G, R, P = p.n_seg, p.n_fsr, p.n_group, p.n_ray, p.n_polar
    return 8 * S + 16 * F * G + 16 * R * P * G


def progress(msg: str):
    """One line per finished section on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def read_pmc(path: str) -> dict:
    """Key: value lines of the committed rocprofv3 PMC summary of the sweep
    kernel, per inner sweep (HBM bytes from FETCH_SIZE + WRITE_SIZE calibrated
    on a copy kernel; VALU instructions from SQ_INSTS_VALU)."""
    vals = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    try:
                        vals[k.strip()] = float(v.strip())
                    except ValueError:
                        pass
    return vals


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hp, n_sweeps: int, n_sweeps_1t: int):
    """FP64 oracle (reference loop structure, OpenMP) timed on this host:
    all host threads (the baseline value) and one thread."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1

    def median_ms(n_threads, n):
        o = Oracle(hp, n_inner=1, gs_boundary=True, energy_mode=0, n_threads=n_threads)
        o.initialize()
        o.calc_fission_source(1.0)
        for g in range(hp.n_group):
            o.build_source(g)
        return o.time_sweeps(n)

    U = hp.updates_per_sweep()
    ms = median_ms(threads, n_sweeps)
    ms1 = median_ms(1, n_sweeps_1t)
    return {"value": U / (ms * 1e-3), "unit": "updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "single_thread": {"value": U / (ms1 * 1e-3), "cores": 1, "ms_per_sweep": ms1,
                              "sample": f"median of {n_sweeps_1t} sweeps after 1 warm-up"},
            "ms_per_sweep": ms,
            "sample": f"median of {n_sweeps} all-group inner sweeps of the same 17x17/7-group "
                      f"deck (after 1 warm-up), FP64, sweep1g only, OMP threads={threads}"}


K_TOL, PSI_TOL, MAX_IT = 1e-7, 1e-6, 3000


def oracle_k(factory, case):
    """FP64 eigenvalue of a branch case from the oracle, run with the
    device's iteration scheme (inner iterations, CMFD, energy mode): Jacobi
    and Gauss-Seidel in energy stop at different distances from the fixed
    point (up to 2e-6 in k on 7-group pins at k_tol 1e-7)."""
    from oracle.binding import Oracle
    hp = factory.host_problem(case)
    o = Oracle(hp, n_inner=factory.n_inner, use_cmfd=factory.use_cmfd, energy_mode=factory.energy_mode)
    rc, st, _ = o.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
    return st.k


def bench_batched(work_dir: str, n_cases: int, n_serial: int, n_oracle: int, use_cmfd: bool = True, lanes: int = 2):
    """Batched branch cases (one graph-replayed batch) against the same
    problems solved one at a time with the single-problem device path; keff
    of a sample checked against the FP64 oracle. With CMFD: 5 inner
    iterations per outer, as mocc's CMFD decks use; without: 1."""
    from mocc_amd.branch import BranchFactory, c4_cases
    from mocc_amd.sweeper import GpuBatchSet
    n_inner = 5 if use_cmfd else 1
    factory = BranchFactory(os.path.join(work_dir, "batch_cmfd" if use_cmfd else "batch"), n_inner=n_inner,
                            use_cmfd=use_cmfd)
    cases = c4_cases(n_cases)
    # the band layout of the whole batch, timed on a sample spread over the mix
    t0 = time.perf_counter()
    factory.tune(cases[::max(1, n_cases // 32)])
    t_tune = time.perf_counter() - t0
    tuned = {"band_rays": factory.band_rays, "lane_map": factory.lane_map,
             "candidates_s": {f"{w}/{m}": round(t, 4) for w, m, t in factory.tuning}, "tune_s": t_tune}
    t0 = time.perf_counter()
    handles = [factory.make(c) for c in cases]
    t_setup = time.perf_counter() - t0
    batch = GpuBatchSet(lanes, use_graph=True)
    ids = [batch.add(h, K_TOL, PSI_TOL, MAX_IT) for h in handles]
    t0 = time.perf_counter()
    res = batch.solve_all(chunk=8)
    t_batch = time.perf_counter() - t0
    progress(f"batched solve of {n_cases} cases: {t_batch:.2f} s")
    k_batch = {i: res[cid].k for i, cid in enumerate(ids)}
    iters = [res[cid].iterations for cid in ids]
    n_conv = sum(res[cid].converged for cid in ids)
    batch.close()
    for h in handles:
        h.close()
    # serial: every (n_cases // n_serial)-th case, same mix
    step = max(1, n_cases // n_serial)
    sub = list(range(0, n_cases, step))[:n_serial]
    sh = [factory.make(cases[i]) for i in sub]
    t0 = time.perf_counter()
    k_serial = {}
    for i, h in zip(sub, sh):
        rc, st = h.eigen_solve(K_TOL, PSI_TOL, MAX_IT)
        k_serial[i] = st.k
    t_serial = time.perf_counter() - t0
    progress(f"serial solve of {len(sub)} cases: {t_serial:.2f} s")
    for h in sh:
        h.close()
    # keff against the FP64 oracle on cases spread over the whole mix
    # (single pins, 3x3 lattices and 17x17 assemblies)
    sample = list(range(0, n_cases, max(1, n_cases // max(n_oracle, 1))))[:n_oracle]
    dk = []
    for i in sample:
        dk.append(abs(k_batch[i] - oracle_k(factory, cases[i])))
        progress(f"oracle keff of case {i} ({cases[i].kind}, {cases[i].groups}g, {cases[i].spacing} cm): "
                 f"|dk| = {dk[-1]:.2e}")
    factory.close()
    bat = n_cases / t_batch
    ser = len(sub) / t_serial
    return {
        "workload": f"{n_cases} branch cases: 1/4 single pin, 1/2 3x3, 1/4 17x17 (25 guide tubes); "
                    "2/7 groups alternating; CY 16x3; spacing from {0.02, 0.035, 0.05} cm; XS x U[0.9,1.1]; "
                    "band width and lane map autotuned for the batch",
        "batch_lanes": lanes,
        "batched_solves_per_s": bat, "serial_solves_per_s": ser, "speedup": bat / ser,
        "serial_sample": len(sub), "converged": int(n_conv), "outer_iterations_max": int(max(iters)),
        "batch_s": t_batch, "setup_s": t_setup,
        "keff_max_abs_err_vs_fp64": max(dk) if dk else None, "keff_checked": len(dk),
        "keff_checked_kinds": sorted({cases[i].kind for i in sample}),
        "autotuned": tuned,
        "keff_max_abs_diff_batch_vs_serial": max(abs(k_batch[i] - k_serial[i]) for i in sub),
        "solver": (f"{'CMFD, ' if use_cmfd else 'no CMFD, '}{n_inner} inner iteration(s) per outer, "
                   "Jacobi in energy, Gauss-Seidel boundary update, k_tol 1e-7, psi_tol 1e-6"),
    }


def bench_serving(work_dir: str, n_req: int, rates=(0.5, 1.0), seed: int = 5, use_cmfd: bool = True,
                  n_serial: int = 64, n_oracle: int = 8, lanes: int = 1):
    """Poisson arrivals of branch-case requests served by the continuous-
    batching loop at the given fractions of its burst (saturation)
    throughput, beside the same requests solved one at a time (CMFD with 5
    inner iterations per outer, or no CMFD with 1)."""
    import numpy as np
    from mocc_amd.branch import BranchFactory, c5_request
    from mocc_amd.serve import Server, poisson_arrivals, serve_serial, summarize
    factory = BranchFactory(os.path.join(work_dir, "serve_cmfd" if use_cmfd else "serve"), n_inner=5 if use_cmfd else 1,
                            use_cmfd=use_cmfd)
    rng = np.random.default_rng(seed)
    cases = [c5_request(rng, i) for i in range(n_req)]
    for c in cases:  # trace every geometry before timing
        factory.host_problem(c)
    factory.tune(cases[:32])
    srv = Server(factory, chunk=1, lanes=lanes)
    burst = summarize(srv.serve(np.zeros(n_req), cases))
    out = {"workload": f"{n_req} requests drawn from {{pin, 3x3, 17x17, 2x2 of 17x17}} x {{2,7}} groups, "
                       "CY 16x3, spacing 0.05 cm, XS x U[0.9,1.1]; "
                       + ("CMFD, 5 inner iterations" if use_cmfd else "no CMFD, 1 inner iteration")
                       + "; band width and lane map autotuned for the batch on the first 32 requests",
           "tuned": {"band_rays": factory.band_rays, "lane_map": factory.lane_map},
           "batch_lanes": lanes, "burst": burst}
    progress(f"serving burst: {burst['solves_per_s']:.3g} solves/s")
    last = None
    for frac in rates:
        rate = frac * burst["solves_per_s"]
        last = srv.serve(poisson_arrivals(n_req, rate, seed + 1), cases)
        out[f"poisson_{frac:g}x"] = {"rate_per_s": rate, **summarize(last)}
        progress(f"serving at {frac:g}x burst done")
    srv.close()
    if n_oracle > 0 and last is not None:
        # served keffs against the FP64 oracle (same iteration scheme), one
        # request of each lattice kind and group count met
        seen, dk = set(), []
        for r in last:
            key = (r.case.kind, r.case.groups)
            if key in seen:
                continue
            seen.add(key)
            dk.append(abs(r.k - oracle_k(factory, r.case)))
            progress(f"served request {r.index} ({r.case.kind}, {r.case.groups}g): |dk| vs FP64 = {dk[-1]:.2e}")
            if len(dk) >= n_oracle:
                break
        out["keff_max_abs_err_vs_fp64"] = max(dk)
        out["keff_checked"] = len(dk)
    n_ser = min(n_req, n_serial)
    out["serial_burst"] = {"sample": n_ser, **summarize(serve_serial(factory, np.zeros(n_ser), cases[:n_ser]))}
    factory.close()
    return out


def bench_core51(work_dir: str, with_oracle: bool):
    """The 3x3-assembly core (51x51 pins, vacuum east/south, CMFD, 5 inner
    iterations, CY 64x3 at 0.02 cm): full eigenvalue solve on the device with
    the FP64 CMFD; optionally the FP64 oracle's solve of the same deck."""
    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    from mocc_amd.sweeper import GpuMoCSweeper
    t0 = time.perf_counter()
    hp = HostProblem.load(lg.write_config("core51", work_dir))
    t_trace = time.perf_counter() - t0
    par = hp.params()
    t0 = time.perf_counter()
    dev = GpuMoCSweeper(hp)
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc, st = dev.eigen_solve(par.k_tol, par.psi_tol, par.max_iter)
    t_solve = time.perf_counter() - t0
    out = {"workload": "3x3 assemblies of 17x17 (51x51 pins, 3 material sets), 7 groups, CY 64x3, "
                       "spacing 0.02 cm, vacuum east/south, CMFD, 5 inner iterations",
           "n_fsr": hp.n_fsr, "n_seg": hp.p.n_seg, "k": st.k, "outer_iterations": st.iterations,
           "converged": bool(st.converged), "solve_s": t_solve, "trace_s": t_trace, "plan_upload_s": t_plan,
           "updates_per_s": dev.updates_per_sweep() * par.n_inner * st.iterations / t_solve,
           "last_cmfd_power_iterations": dev.cmfd_iterations(),
           "last_cmfd_bicgstab_iterations": dev.cmfd_bicgstab_iterations(),
           "k_tol": par.k_tol, "psi_tol": par.psi_tol}
    dev.close()
    # the FP64 oracle's solve of the same deck, stored with the tests (it
    # takes many minutes on host cores): tests/golden/make_core51_oracle.py
    ref_path = os.path.join(ROOT, "tests", "golden", "core51_oracle.npz")
    if os.path.exists(ref_path):
        import numpy as np
        ref = np.load(ref_path)
        out["oracle"] = {"k": float(ref["k"]), "outer_iterations": int(ref["iterations"]),
                         "keff_abs_err": abs(st.k - float(ref["k"])),
                         "source": "tests/golden/core51_oracle.npz (FP64 oracle, same deck and iteration scheme)"}
    if with_oracle:
        from oracle.binding import Oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        o = Oracle(hp, n_inner=par.n_inner, use_cmfd=True, n_threads=threads)
        t0 = time.perf_counter()
        rc_o, st_o, _ = o.eigen_solve(par.k_tol, par.psi_tol, par.max_iter)
        out["oracle_live"] = {"k": st_o.k, "outer_iterations": st_o.iterations, "solve_s": time.perf_counter() - t0,
                              "cores": threads, "keff_abs_err": abs(st.k - st_o.k)}
    return out


def copy_calibration():
    """A plain device copy of known size: 1 GiB read and 1 GiB written per
    copy, three copies. Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it gives
    the counters' KiB per GiB actually moved, which scales the sweep kernel's
    readings into HBM bytes."""
    import torch
    n = 256 * 1024 * 1024
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-sweeps", type=int, default=20)
    ap.add_argument("--cpu-sweeps-1t", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched branch-case section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving section")
    ap.add_argument("--batch-cases", type=int, default=256)
    ap.add_argument("--serial-cases", type=int, default=32)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--lane-map", type=int, default=-1,
                    help="lane map of the 17x17 sweep: -1 autotunes band width and lane map at create")
    ap.add_argument("--band-rays", type=int, default=0, help="band width of the 17x17 sweep with a fixed lane map")
    ap.add_argument("--batch-lanes", type=int, default=2,
                    help="batches (each on its own stream) the batched branch cases are spread over")
    ap.add_argument("--serve-lanes", type=int, default=1,
                    help="batches (each on its own stream) the serving loop spreads its requests over")
    ap.add_argument("--serve-rates", default="0.5,1.0",
                    help="comma-separated arrival rates as fractions of the burst throughput")
    ap.add_argument("--no-core51", action="store_true",
                    help="skip the full eigen solve of the 51x51-pin core (3x3 assemblies, 7 groups, CMFD)")
    ap.add_argument("--core51-oracle", action="store_true",
                    help="also run the FP64 oracle's solve of the 51x51-pin core live (many minutes)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs on one GPU only "
              "(multi-GPU sharding is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration()
        return

    from mocc_amd import lattice_gen as lg
    from mocc_amd.native import HostProblem
    from mocc_amd.sweeper import GpuMoCSweeper

    deck_dir = tempfile.mkdtemp(prefix="mocc_bench_")
    path = lg.write_config("assembly17", deck_dir)
    hp = HostProblem.load(path)
    p = hp.p
    # band width and lane map picked at create by timing the candidates
    # (--lane-map -1), or as given
    t0 = time.perf_counter()
    dev = GpuMoCSweeper(hp, n_inner=1, energy_mode=1, lane_map=args.lane_map, band_rays=args.band_rays)
    t_tune = time.perf_counter() - t0
    tuned = dev.options()
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
    pmc = read_pmc(os.path.join(ROOT, "profiles", "sweep_assembly17_pmc.txt"))
    # the counters were taken with one band layout; they stand for this run
    # only when it runs the same one
    if (pmc.get("band_rays"), pmc.get("lane_map")) != (float(tuned.band_rays), float(tuned.lane_map)):
        pmc = {}
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": pmc.get("hbm_bytes_per_sweep"),
                "kernel": f"k_sweep<3,false>, both Gauss-Seidel phase launches of one sweep "
                          f"({int(tuned.band_rays)}-ray bands, lane map {int(tuned.lane_map)})",
                "kernel_ms": kernel_ms, "algorithmic_bytes": B}
    insts = pmc.get("sq_insts_valu_per_sweep")
    if insts:
        # the instruction-issue bound that actually limits this kernel
        issue_ms = insts * VALU_CYCLES_PER_INST / N_SIMD / CLOCK_HZ * 1e3
        roofline["valu_issue"] = {"insts_per_sweep": insts, "issue_ms": issue_ms,
                                  "frac": issue_ms / kernel_ms}
    line = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "17x17 assembly (25 guide tubes), 7 groups, CY 32x3, spacing 0.03 cm, "
                               "reflective; one all-group inner sweep per step",
                   "n_fsr": p.n_fsr, "n_seg": p.n_seg, "n_ray": p.n_ray, "updates_per_step": U,
                   "band_rays": int(tuned.band_rays), "lane_map": int(tuned.lane_map),
                   "autotune_s": round(t_tune, 3), "parallelism": "single GPU"},
        "roofline": roofline,
    }
    progress(f"17x17 7-group sweep: {value:.4g} updates/s")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(hp, args.cpu_sweeps, args.cpu_sweeps_1t)
        progress("CPU baseline done")
    dev.close()
    if not args.no_batch:
        line["batched"] = bench_batched(deck_dir, args.batch_cases, args.serial_cases, 0 if args.no_cpu else 8,
                                        lanes=args.batch_lanes)
        progress(f"batched (CMFD): {line['batched']['batched_solves_per_s']:.4g} solves/s")
        line["batched_no_cmfd"] = bench_batched(deck_dir, args.batch_cases, args.serial_cases,
                                                0 if args.no_cpu else 2, use_cmfd=False, lanes=args.batch_lanes)
        progress(f"batched (no CMFD): {line['batched_no_cmfd']['batched_solves_per_s']:.4g} solves/s")
    if not args.no_serve:
        rates = tuple(float(x) for x in args.serve_rates.split(",") if x)
        line["serving"] = bench_serving(deck_dir, args.serve_requests, rates, n_oracle=0 if args.no_cpu else 8,
                                        lanes=args.serve_lanes)
        progress("serving done")
    if not args.no_core51:
        line["core51"] = bench_core51(deck_dir, args.core51_oracle)
        progress("51x51-pin core solve done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
