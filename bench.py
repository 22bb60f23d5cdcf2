"""Benchmark of the nonlinear FVM step on one MI355X.

Workload: 2D2V Maxwellian perturbation, fields off, N_q1 = N_q2 = 128 periodic
(3 ghost zones in the reference's layout), N_p1 = N_p2 = 48 on [-10, 10]^2
(37,748,736 phase-space cells), WENO5 reconstruction, LLF flux, BGK tau = 1e-2,
SSP-RK3, dt = 5e-4, plain FP64 on the GPU.  A step is one full SSP-RK3 step
(three fused stage kernels).  The FP32-bulk form of the same step is timed
beside it and its rel-L2 against the FP64 step is reported.

Prints one JSON line.  ``value`` is FP64 phase-space cell-updates/s with f
resident in HBM.  ``roofline`` is for the dominant kernel (the fused stage),
with the device-copy bandwidth measured on the same GPU beside the nominal
peak (``copy_gbs``, ``frac_of_copy``):
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


def copy_bandwidth(mib=1024, reps=10):
    """Achievable device-copy bandwidth on this GPU: torch's copy of a
    ``mib`` MiB buffer, read + write bytes over the mean time of ``reps``
    copies (HIP events)."""
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    s = e0.elapsed_time(e1) / 1e3 / reps
    return 2 * n * 8 / s / 1e9


def run_adjoint(steps=100, every=10):
    """The adjoint workload: 2D2V 64^2 x 32^2, WENO5 + LLF + BGK tau = 0.05,
    SSP-RK3 dt = 5e-4, 100 steps with checkpoints every 10.  Times the
    forward run and the backward run (segment recompute + stage adjoints)."""
    import torch
    from bolt_amd import adjoint as A
    from tests.cases import case_c4, hip_solver, maxwellian_interior
    c = case_c4(seed=0)
    s = hip_solver(c)
    v = [x.reshape(-1) for x in (s.p1_center, s.p2_center, s.p3_center)]
    dq = s.dq1 * s.dq2
    dn = c.dn.to(s.device).requires_grad_(True)
    tau = torch.tensor(0.05, dtype=torch.float64, device=s.device, requires_grad=True)

    def once():
        fT = A.evolve(s, maxwellian_interior(*v, 2, dn), tau, steps, c.dt, every)
        n, E = A.moment(s, fT, "density"), A.moment(s, fT, "energy")
        J = ((n - 1) ** 2).sum() * dq + E.sum() * dq
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        torch.autograd.grad(J, [dn, tau])
        torch.cuda.synchronize()
        return t1, time.perf_counter()

    once()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t1, t2 = once()
    cells = s._Q * s._P
    return {"workload": "2D2V adjoint 64^2 q x 32^2 p, WENO5 + LLF + BGK tau=0.05, SSP-RK3 "
                        "dt=5e-4, %d steps, checkpoints every %d, FP64" % (steps, every),
            "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3,
            "cell_steps_per_s_forward_plus_backward": cells * steps / (t2 - t0)}


def run_em(steps=10, warmup=2):
    """2D3V electromagnetic step: 64^2 q x 24^3 p on [-8, 8]^3, electrons
    (q/m = -1) in B = (0, 0, 0.5), PPM in q, WENO5 in p, LLF, BGK tau = 0.1,
    SSP-RK3 dt = 2e-4, Yee FDTD fields; FP32-bulk and plain FP64, with the
    FP32-bulk error against the FP64 run after the same steps."""
    import torch
    from tests.cases import case_2d3v, hip_solver
    out = {}
    f = {}
    for prec in ("fp64", "fp32-bulk"):
        c = case_2d3v(seed=0, precision=prec)
        s = hip_solver(c)
        for _ in range(warmup):
            s._advance(c.dt)
        events = []
        s._stage_events = events
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s._advance(c.dt)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        s._stage_events = None
        s.check_divergence()
        # the stage kernel (qcell_kernel with p-advection) against the HBM
        # roofline at the algorithmic bytes of an SSP-RK3 stage (8/3 f-passes)
        stage_s = sum(e0.elapsed_time(e1) for e0, e1 in events) / len(events) / 1e3
        esz = 8 if prec == "fp64" else 4
        achieved = 8 / 3 * s._Q * s._P * esz / stage_s / 1e9
        out[prec] = {"ms_per_step": dt * 1e3, "value": s._Q * s._P / dt,
                     "stage_launch_ms_avg": stage_s * 1e3,
                     "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                  "traffic": None}}
        f[prec] = s._f.double()
        del s
    rel = (torch.linalg.vector_norm(f["fp32-bulk"] - f["fp64"]) /
           torch.linalg.vector_norm(f["fp64"])).item()
    return {"workload": "2D3V EM 64^2 q x 24^3 p, PPM q / WENO5 p, LLF, BGK tau=0.1, "
                        "SSP-RK3 dt=2e-4, FDTD fields",
            "cells": 64 * 64 * 24 ** 3, "fp32_bulk": out["fp32-bulk"], "fp64": out["fp64"],
            "fp32_bulk_rel_l2_vs_fp64_after_%d_steps" % (steps + warmup): rel,
            "speedup_fp32_bulk_vs_fp64": out["fp32-bulk"]["value"] / out["fp64"]["value"]}


def run_ensemble(steps=200, sample_steps=10):
    """The ensemble workload: 96 1D2V members of uneven sizes (N_q1 in
    {32..256}, N_p in {24..64}^2, tau in [1e-3, 1], own dt each), WENO5 + LLF
    + BGK, SSP-RK3, FP32-bulk with FP64 moments, 200 steps in persistent
    ensemble launches with the autotuned work-list shape, replayed as one HIP
    graph (the eager time is reported too); beside it the same
    members stepped one after another by the single-problem solver in plain
    FP64 (timed over ``sample_steps`` steps each)."""
    import torch
    from bolt_amd.ensemble import Ensemble
    from tests.cases import case_c5, hip_solver
    m32 = case_c5(seed=0, precision="fp32-bulk")
    m64 = case_c5(seed=0, precision="fp64")
    dts = [dt for _, dt in m32]
    t0 = time.perf_counter()
    ens = Ensemble([hip_solver(c) for c, _ in m32], dts)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    cells = sum(s._Q * s._P for s in ens.solvers)
    ens.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    # the same steps as one HIP graph (the launches recorded once, replayed)
    ens.capture(steps)
    ens.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.replay()
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - t0
    ens.check_divergence()
    # sequential FP64 baseline and the FP32-bulk error against it
    solo = [hip_solver(c) for c, _ in m64]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, (c, dt) in zip(solo, m64):
        for _ in range(sample_steps):
            s.strang_timestep(dt)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    chk = Ensemble([hip_solver(c) for c, _ in m32], dts, tuning=ens.tuning)
    chk.step(sample_steps)
    worst = 0.0
    for i, s in enumerate(solo):
        a, b = chk.state(i).double(), chk._lines(s)
        worst = max(worst, float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
    rate = cells * steps / t_ens
    seq = cells * sample_steps / t_seq
    return {"workload": "96 x 1D2V ensemble (N_q1 in {32..256}, N_p in {24..64}^2, own dt and "
                        "tau), WENO5 + LLF + BGK, SSP-RK3, FP32-bulk, %d steps" % steps,
            "members": len(dts), "cells": cells, "value": rate,
            "unit": "phase-space cell-updates/s", "ms_per_step": t_ens / steps * 1e3,
            "ms_per_step_without_graph": t_eager / steps * 1e3,
            "tuning": {"waves_per_group": ens.tuning[0], "cells_per_task": ens.tuning[1],
                       "groups_per_cu": ens.tuning[2], "chunk_cells": ens.tuning[3]},
            "autotune_s": t_tune,
            "sequential_fp64": {"value": seq, "sample": "%d steps per member, one member "
                                                        "after another" % sample_steps},
            "speedup_vs_sequential_fp64": rate / seq,
            "max_rel_l2_fp32_bulk_vs_fp64_after_%d_steps" % sample_steps: worst}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The FP64 oracle (torch CPU eager, the reference's whole-array op
    sequence) on bounded samples of the same problem: 64^2 q-cells x 48^2
    p-cells with every host thread this process may use, and 32^2 x 24^2 on
    one thread; one SSP-RK3 step each."""
    import torch
    from tests.cases import case_2d2v, oracle_solver
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1

    def rate(threads, nq, npp):
        torch.set_num_threads(threads)
        c = case_2d2v(seed=0, N_q=nq, N_p=npp)
        o = oracle_solver(c)
        t0 = time.perf_counter()
        o.strang_timestep(c.dt)
        return nq * nq * npp * npp / (time.perf_counter() - t0)

    v = rate(n, 64, 48)
    v1 = rate(1, 32, 24)
    torch.set_num_threads(n)
    return {"value": v, "unit": "phase-space cell-updates/s", "cores": n,
            "kind": "port",
            "sample": "FP64 oracle (torch CPU), 2D2V WENO5/LLF/BGK SSP-RK3 at "
                      "64^2 x 48^2 cells, 1 step, %d threads" % n,
            "cpu_model": _cpu_model(),
            "one_thread": {"value": v1, "sample": "32^2 x 24^2 cells, 1 step, 1 thread"}}


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
    ap.add_argument("--steps", type=int, default=100)
    # the stage kernel reaches its steady rate after about 30 steps (clocks
    # ramp up from idle): warm up past that
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 2D3V electromagnetic and adjoint measurements")
    ap.add_argument("--probe", choices=["fp64", "fp32-bulk"],
                    help="traffic probe for rocprofv3 --pmc: a 256 MiB device copy "
                         "(the counter calibration) and then one step of the workload")
    ap.add_argument("--only", choices=["em_2d3v", "adjoint", "ensemble"],
                    help="run one of the other workloads alone (for profiling) and print "
                         "its JSON object")
    args = ap.parse_args()
    if args.probe:
        probe(args.probe)
        return
    if args.only:
        run = {"em_2d3v": run_em, "adjoint": run_adjoint, "ensemble": run_ensemble}[args.only]
        print(json.dumps({args.only: run()}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the step shards over q with a halo exchange per stage)", file=sys.stderr)
        sys.exit(2)
    import torch
    copy_gbs = copy_bandwidth()
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
        "roofline": dict(r64["roofline"], copy_gbs=copy_gbs,
                         frac_of_copy=r64["roofline"]["achieved"] / copy_gbs),
        "fp32_bulk": {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                      "dtype": "f32 bulk, f64 moments/temperature",
                      "roofline": dict(r32["roofline"], copy_gbs=copy_gbs,
                                       frac_of_copy=r32["roofline"]["achieved"] / copy_gbs),
                      "rel_l2_vs_fp64_after_run": rel_l2,
                      "speedup_vs_fp64": r32["value"] / r64["value"]},
    }
    if not args.no_extras:
        out["em_2d3v"] = run_em()
        out["adjoint"] = run_adjoint()
        out["ensemble"] = run_ensemble()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
