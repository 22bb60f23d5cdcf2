"""Benchmark of the nonlinear FVM step on one MI355X.

Workload: 2D2V Maxwellian perturbation, fields off, N_q1 = N_q2 = 128 periodic
(3 ghost zones in the reference's layout), N_p1 = N_p2 = 48 on [-10, 10]^2
(37,748,736 phase-space cells), WENO5 reconstruction, LLF flux, BGK tau = 1e-2,
SSP-RK3, dt = 5e-4, plain FP64 on the GPU.  A step is one full SSP-RK3 step
(three fused stage kernels).  The FP32-bulk form of the same step is timed
beside it and its rel-L2 against the FP64 step is reported.

Prints one JSON line.  ``value`` is FP64 phase-space cell-updates/s with f
resident in HBM.  ``roofline`` is for the dominant kernel (the fused stage):
algorithmic bytes (each stage charged one read per f-sized input and one write,
8 passes per SSP-RK3 step = 64 B/cell in FP64, 32 B/cell in FP32-bulk) divided
by its measured launch time (HIP events on the launch stream).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("phase-space cell-updates/s, FP32-bulk vs FP64 at rel-L2≤1e-5; "
          "% HBM peak")


def run_gpu(precision, steps, warmup):
    import torch
    from tests.cases import case_2d2v, hip_solver
    c = case_2d2v(seed=0, precision=precision)
    nls = hip_solver(c)
    for _ in range(warmup):
        nls._advance(c.dt)
    nls.check_divergence()
    events = []
    nls._stage_events = events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nls._advance(c.dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nls.check_divergence()
    nls._stage_events = None
    stage_ms = [e0.elapsed_time(e1) for e0, e1 in events]
    cells = nls._Q * nls._P
    bytes_per_cell_step = 8 * (8 if precision == 'fp64' else 4)
    kernel_s = sum(stage_ms) / 1e3
    achieved = bytes_per_cell_step * cells * steps / kernel_s / 1e9
    traffic = _traffic(precision)
    return nls, {
        "cells": cells,
        "ms_per_step": wall / steps * 1e3,
        "value": cells * steps / wall,
        "stage_launch_ms_avg": sum(stage_ms) / len(stage_ms),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernel": "qline_kernel (stage stencil; SSP-RK3 stages average 8/3 f-passes)",
    }


def _traffic(precision):
    """HBM bytes per stage-kernel launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration); the
    measured figures are kept in profiles/traffic.json."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(precision, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline():
    """The FP64 oracle (torch CPU, all host threads) on a bounded sample of the
    same problem: 64^2 q-cells x 48^2 p-cells, one SSP-RK3 step."""
    import torch
    from tests.cases import case_2d2v, oracle_solver
    n = os.cpu_count() or 1
    torch.set_num_threads(n)
    c = case_2d2v(seed=0, N_q=64)
    o = oracle_solver(c)
    t0 = time.perf_counter()
    o.strang_timestep(c.dt)
    dt = time.perf_counter() - t0
    cells = 64 * 64 * 48 * 48
    return {"value": cells / dt, "unit": "phase-space cell-updates/s", "cores": n,
            "kind": "port",
            "sample": "FP64 oracle (torch CPU), 2D2V WENO5/LLF/BGK SSP-RK3 at "
                      "64^2 x 48^2 cells, 1 step (%.1f s)" % dt}


def probe(precision):
    """Launch sequence for counter runs: three copies of a known size, then one
    SSP-RK3 step (three stage kernels and their moment passes)."""
    import torch
    from tests.cases import case_2d2v, hip_solver
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    c = case_2d2v(seed=0, precision=precision)
    nls = hip_solver(c)
    nls._advance(c.dt)
    torch.cuda.synchronize()
    print("probe done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", choices=["fp64", "fp32-bulk"],
                    help="traffic probe for rocprofv3 --pmc: a 256 MiB device copy "
                         "(the counter calibration) and then one step of the workload")
    args = ap.parse_args()
    if args.probe:
        probe(args.probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the step shards over q with a halo exchange per stage)", file=sys.stderr)
        sys.exit(2)
    import torch
    nls64, r64 = run_gpu('fp64', args.steps, args.warmup)
    nls32, r32 = run_gpu('fp32-bulk', args.steps, args.warmup)
    # accuracy of the FP32-bulk step against the FP64 step, same inputs, same steps
    f64 = nls64._f
    f32 = nls32._f.double()
    rel_l2 = (torch.linalg.vector_norm(f32 - f64) / torch.linalg.vector_norm(f64)).item()
    del nls64, nls32
    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "phase-space cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Maxwellian density perturbation)",
        "config": {"workload": "2D2V FVM fields off, 128^2 q x 48^2 p, WENO5 + LLF + "
                               "BGK tau=1e-2, SSP-RK3 dt=5e-4, plain FP64",
                   "cells": r64["cells"], "parallelism": "single GPU"},
        "roofline": r64["roofline"],
        "fp32_bulk": {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                      "dtype": "f32 bulk, f64 moments/temperature",
                      "roofline": r32["roofline"],
                      "rel_l2_vs_fp64_after_run": rel_l2,
                      "speedup_vs_fp64": r32["value"] / r64["value"]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
