"""Benchmark of the nonlinear FVM step on one MI355X.

Workload: 2D2V Maxwellian perturbation, fields off, N_q1 = N_q2 = 128 periodic
(3 ghost zones in the reference's layout), N_p1 = N_p2 = 48 on [-10, 10]^2
(37,748,736 phase-space cells), WENO5 reconstruction, LLF flux, BGK tau = 1e-2,
SSP-RK3, dt = 5e-4, plain FP64 on the GPU.  A step is one full SSP-RK3 step
(three fused stage kernels).  The FP32-bulk form of the same step is timed
beside it and its rel-L2 against the FP64 step is reported.

Prints one JSON line.  ``value`` is FP64 phase-space cell-updates/s with f
resident in HBM: after at least 1 s of warm-up steps (``--warmup`` steps or
more), exactly ``--steps`` steps are timed in five repeats of near-equal length
and the median repeat gives ``ms_per_step``.  ``roofline`` is for the dominant
kernel (the fused stage), with the device-copy bandwidth measured on the same
GPU beside the nominal peak (``copy_gbs``, ``frac_of_copy``): algorithmic bytes
(each stage charged one read per f-sized input and one write, 8 passes per
SSP-RK3 step = 64 B/cell in FP64, 32 B/cell in FP32-bulk) divided by its
measured launch time (HIP events on the launch stream); ``step_achieved`` /
``step_frac`` divide the same bytes by the whole step's time (every launch of
the step, moment passes included).  ``traffic`` is the counter-measured HBM
bytes per launch (``--probe`` / ``--traffic``).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# work shape of the ensemble traffic probe (the autotuner's pick for the
# 96-member workload on MI355X)
ENSEMBLE_PROBE_TUNING = (32768, 2048)
METRIC = ("phase-space cell-updates/s, FP32-bulk vs FP64 at rel-L2≤1e-5; "
          "% HBM peak")


def warm_up(step, warmup, min_seconds=1.0):
    """``warmup`` untimed steps, then more until ``min_seconds`` of stepping
    have passed (the clocks ramp up from idle over the first ~0.5 s, whatever
    step count the caller asked for).  Returns the steps run."""
    import torch
    t0 = time.perf_counter()
    n = 0
    while n < warmup or time.perf_counter() - t0 < min_seconds:
        step()
        n += 1
        if n >= warmup:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    return n


def timed_repeats(step, steps, repeats=5):
    """Exactly ``steps`` steps, split into ``min(repeats, steps)`` timed
    repeats of near-equal length, each bracketed by device synchronisation.
    Returns the ms per step of every repeat and the total wall seconds."""
    import torch
    r = max(1, min(repeats, steps))
    sizes = [steps // r + (1 if k < steps % r else 0) for k in range(r)]
    ms, total = [], 0.0
    for n in sizes:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        total += dt
        ms.append(dt / n * 1e3)
    return ms, total


def _median(x):
    y = sorted(x)
    n = len(y)
    return y[n // 2] if n % 2 else 0.5 * (y[n // 2 - 1] + y[n // 2])


def run_gpu(precision, steps, warmup):
    from tests.cases import case_2d2v, hip_solver
    c = case_2d2v(seed=0, precision=precision)
    nls = hip_solver(c)
    warm = warm_up(lambda: nls._advance(c.dt), warmup)
    nls.check_divergence()
    events = []
    nls._stage_events = events
    ms_eager, _ = timed_repeats(lambda: nls._advance(c.dt), steps)
    nls.check_divergence()
    nls._stage_events = None
    stage_ms = [e0.elapsed_time(e1) for e0, e1 in events]
    # the timed steps: a one-step HIP graph of the same launches, replayed
    # exactly ``steps`` times (no launch gaps between the step's kernels)
    nls.capture(1, c.dt)
    warm_up(nls.replay, 5, min_seconds=0.2)
    ms, wall = timed_repeats(nls.replay, steps)
    nls.check_divergence()
    cells = nls._Q * nls._P
    bytes_per_cell_step = 8 * (8 if precision == 'fp64' else 4)
    kernel_s = sum(stage_ms) / 1e3
    achieved = bytes_per_cell_step * cells * steps / kernel_s / 1e9
    ms_step = _median(ms)
    step_gbs = bytes_per_cell_step * cells / (ms_step / 1e3) / 1e9
    traffic = _traffic(precision)
    return nls, {
        "cells": cells,
        "ms_per_step": ms_step,
        "ms_per_step_repeats": ms,
        "ms_per_step_without_graph": _median(ms_eager),
        "warmup_steps_run": warm,
        "timed_wall_s": wall,
        "value": cells / (ms_step / 1e3),
        "stage_launch_ms_avg": sum(stage_ms) / len(stage_ms),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "step_achieved": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS},
        "kernel": "qline_kernel (stage stencil; SSP-RK3 stages average 8/3 f-passes)",
    }


def _traffic(key, field="bytes_per_launch"):
    """HBM bytes per launch of the workload's dominant kernel (or, with
    ``field="bytes_per_sequence"``, summed over the launches of one step)
    measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy-kernel calibration), as ``bench.py --traffic`` wrote them into
    profiles/traffic.json; None if not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(key, {}).get(field)
    except (OSError, ValueError):
        return None


def _pmc_rows(path, counter):
    """(kernel name, value) per dispatch of ``counter`` from a rocprofv3
    counter-collection CSV, in dispatch order."""
    import csv
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if r.get("Counter_Name") == counter:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def traffic_from_pmc(key, fetch_csv, write_csv, kernel, algorithmic, after=None, extra=None):
    """HBM bytes per launch of ``kernel`` (names separated by '|') from one
    FETCH_SIZE run and one WRITE_SIZE run of ``bench.py --probe`` (KiB per
    dispatch; with ``after``, only the dispatches from the first one whose
    name holds ``after`` on).  Each counter
    is scaled by what it reads for the probe's first 256 MiB device copy
    (read 256 MiB, write 256 MiB) in the same run.  ``algorithmic`` are the
    algorithmic bytes of each launch, in order.  Updates
    profiles/traffic.json under ``key`` and returns the entry."""
    mib256 = 256 * 1024.0  # KiB
    out = {}
    for name, path in (("fetch", fetch_csv), ("write", write_csv)):
        rows = _pmc_rows(path, name.upper() + "_SIZE")
        cal = [v for k, v in rows if "copy" in k.lower()]
        if after:
            first = next((i for i, (k, _) in enumerate(rows) if after in k), len(rows))
            rows = rows[first:]
        launches = [v for k, v in rows if any(n in k for n in kernel.split("|"))]
        if not cal or not launches:
            raise SystemExit("no copy or %s dispatches in %s" % (kernel, path))
        out[name + "_scale"] = mib256 / cal[0]
        out[name + "_kib_per_launch"] = launches
    n = min(len(out["fetch_kib_per_launch"]), len(out["write_kib_per_launch"]), len(algorithmic))
    per = [1024.0 * (out["fetch_kib_per_launch"][i] * out["fetch_scale"] +
                     out["write_kib_per_launch"][i] * out["write_scale"]) for i in range(n)]
    entry = dict(out, kernel=kernel, per_launch_bytes=per,
                 algorithmic_per_launch_bytes=list(algorithmic[:n]),
                 bytes_per_launch=sum(per) / n, bytes_per_sequence=sum(per),
                 ratio_to_algorithmic=sum(per) / sum(algorithmic[:n]), **(extra or {}))
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        table = {}
    table[key] = entry
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1)
    return entry


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


def run_adjoint(steps=100, every=10, repeats=3):
    """The adjoint workload: 2D2V 64^2 x 32^2, WENO5 + LLF + BGK tau = 0.05,
    SSP-RK3 dt = 5e-4, 100 steps with checkpoints every 10.  Times the
    forward run and the backward run (segment recompute + stage adjoints)
    after one untimed run; the medians of ``repeats`` runs are reported."""
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
    fwd, bwd = [], []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t1, t2 = once()
        fwd.append((t1 - t0) * 1e3)
        bwd.append((t2 - t1) * 1e3)
    cells = s._Q * s._P
    f_ms, b_ms = _median(fwd), _median(bwd)
    return {"workload": "2D2V adjoint 64^2 q x 32^2 p, WENO5 + LLF + BGK tau=0.05, SSP-RK3 "
                        "dt=5e-4, %d steps, checkpoints every %d, FP64" % (steps, every),
            "forward_ms": f_ms, "backward_ms": b_ms,
            "forward_ms_repeats": fwd, "backward_ms_repeats": bwd,
            "backward_over_forward": b_ms / f_ms,
            "cell_steps_per_s_forward_plus_backward": cells * steps / ((f_ms + b_ms) / 1e3)}


def run_em(steps=10, warmup=2):
    """2D3V electromagnetic step: 64^2 q x 24^3 p on [-8, 8]^3, electrons
    (q/m = -1) in B = (0, 0, 0.5), PPM in q, WENO5 in p, LLF, BGK tau = 0.1,
    SSP-RK3 dt = 2e-4 with the fields advanced once per step (Yee FDTD);
    FP32-bulk and plain FP64, with the FP32-bulk error against the FP64 run
    after the same steps.  The roofline is that of the stage kernel
    (qcell_kernel with p-advection) at the algorithmic bytes of an SSP-RK3
    stage (8/3 f-passes), beside the whole step's algorithmic rate."""
    import torch
    from tests.cases import case_2d3v, hip_solver
    out = {}
    for prec in ("fp64", "fp32-bulk"):
        c = case_2d3v(seed=0, precision=prec)
        s = hip_solver(c)
        warm = warm_up(lambda: s._advance(c.dt), warmup)
        events = []
        s._stage_events = events
        ms, _ = timed_repeats(lambda: s._advance(c.dt), steps)
        s._stage_events = None
        s.check_divergence()
        stage_s = sum(e0.elapsed_time(e1) for e0, e1 in events) / len(events) / 1e3
        esz = 8 if prec == "fp64" else 4
        cells = s._Q * s._P
        achieved = 8 / 3 * cells * esz / stage_s / 1e9
        ms_step = _median(ms)
        step_gbs = 8 * esz * cells / (ms_step / 1e3) / 1e9
        out[prec] = {"ms_per_step": ms_step, "ms_per_step_repeats": ms,
                     "warmup_steps_run": warm, "value": cells / (ms_step / 1e3),
                     "stage_launch_ms_avg": stage_s * 1e3,
                     "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                  "traffic": _traffic("em_" + prec),
                                  "step_achieved": step_gbs,
                                  "step_frac": step_gbs / HBM_PEAK_GBS},
                     "kernel": "qcell_kernel (stage with p-advection, 8/3 f-passes per launch)"}
        del s
    n_err = 20
    return {"workload": "2D3V EM 64^2 q x 24^3 p, PPM q / WENO5 p, LLF, BGK tau=0.1, "
                        "SSP-RK3 dt=2e-4, FDTD fields once per step",
            "cells": 64 * 64 * 24 ** 3, "fp32_bulk": out["fp32-bulk"], "fp64": out["fp64"],
            "fp32_bulk_rel_l2_vs_fp64_after_%d_steps" % n_err: fp32_bulk_error(case_2d3v, n_err),
            "speedup_fp32_bulk_vs_fp64": out["fp32-bulk"]["value"] / out["fp64"]["value"]}


def fp32_bulk_error(case_fn, n_steps):
    """rel-L2 of f between the FP32-bulk and the plain FP64 GPU step after
    ``n_steps`` steps from the same initial state (untimed)."""
    import torch
    from tests.cases import hip_solver
    f = {}
    for prec in ("fp64", "fp32-bulk"):
        c = case_fn(seed=0, precision=prec)
        s = hip_solver(c)
        for _ in range(n_steps):
            s._advance(c.dt)
        s.check_divergence()
        f[prec] = s._f.double()
        del s
    return (torch.linalg.vector_norm(f["fp32-bulk"] - f["fp64"]) /
            torch.linalg.vector_norm(f["fp64"])).item()


def run_ensemble(steps=200, sample_steps=10):
    """The ensemble workload: 96 1D2V members of uneven sizes (N_q1 in
    {32..256}, N_p in {24..64}^2, tau in [1e-3, 1], own dt each), WENO5 + LLF
    + BGK, SSP-RK3, FP32-bulk with FP64 moments, 200 steps of the ensemble
    launches (a chunk-stage and a finalize launch per stage for all members)
    at the autotuned work shape, replayed as one HIP graph (the eager time is
    reported too); beside it the same
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
    # whole-step algorithmic bytes (SSP-RK3, FP32-bulk: 8 passes of 4 B)
    step_gbs = 32 * cells * steps / t_ens / 1e9
    return {"workload": "96 x 1D2V ensemble (N_q1 in {32..256}, N_p in {24..64}^2, own dt and "
                        "tau), WENO5 + LLF + BGK, SSP-RK3, FP32-bulk, %d steps" % steps,
            "members": len(dts), "cells": cells, "value": rate,
            "unit": "phase-space cell-updates/s", "ms_per_step": t_ens / steps * 1e3,
            "ms_per_step_without_graph": t_eager / steps * 1e3,
            "roofline": {"bound": "hbm", "level": "step (all launches of a step)",
                         "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": step_gbs / HBM_PEAK_GBS,
                         # counter-measured HBM bytes of one step (stage and
                         # finalize launches) at the probe's work shape
                         "traffic": _traffic("ensemble", "bytes_per_sequence"),
                         "traffic_tuning": list(ENSEMBLE_PROBE_TUNING)},
            "tuning": {"chunk_cells": ens.tuning[0], "batch_cells": ens.tuning[1]},
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


def cpu_baseline(steps=3):
    """The FP64 oracle (torch CPU eager, the reference's whole-array op
    sequence) on the full headline grid, 128^2 q-cells x 48^2 p-cells, over
    ``steps`` SSP-RK3 steps with every host thread this process may use; and
    on one thread over one step of 32^2 x 48^2 (the same p-grid, a quarter
    of the q-rows per axis)."""
    import torch
    from tests.cases import case_2d2v, oracle_solver
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1

    def rate(threads, nq, npp, k):
        torch.set_num_threads(threads)
        c = case_2d2v(seed=0, N_q=nq, N_p=npp)
        o = oracle_solver(c)
        t0 = time.perf_counter()
        for _ in range(k):
            o.strang_timestep(c.dt)
        return nq * nq * npp * npp * k / (time.perf_counter() - t0)

    def rate_cases(cases, k):
        cells, t = 0, 0.0
        for c in cases:
            d = c.domain
            o = oracle_solver(c)
            t0 = time.perf_counter()
            for _ in range(k):
                o.strang_timestep(c.dt)
            t += time.perf_counter() - t0
            cells += d.N_q1 * d.N_q2 * d.N_p1 * d.N_p2 * d.N_p3 * k
        return cells / t

    v = rate(n, 128, 48, steps)
    v1 = rate(1, 32, 48, 1)
    torch.set_num_threads(n)
    # the other workloads on bounded samples, same oracle, every thread
    from tests.cases import case_2d3v, case_c5
    v_em = rate_cases([case_2d3v(seed=0, N_q=12)], 1)
    v_ens = rate_cases([c for c, _ in case_c5(seed=0, precision="fp64")[:8]], 1)
    return {"value": v, "unit": "phase-space cell-updates/s", "cores": n,
            "kind": "port",
            "sample": "FP64 oracle (torch CPU), 2D2V WENO5/LLF/BGK SSP-RK3 on the full "
                      "128^2 x 48^2 grid, %d steps, %d threads" % (steps, n),
            "cpu_model": _cpu_model(),
            "one_thread": {"value": v1, "sample": "32^2 x 48^2 cells, 1 step, 1 thread"},
            "thread_speedup": v / v1,
            "em_2d3v": {"value": v_em, "sample": "2D3V EM (PPM q, WENO5 p, FDTD) on 12^2 x 24^3, "
                                                 "1 step, %d threads" % n},
            "ensemble": {"value": v_ens, "sample": "the first 8 of the 96 1D2V members, 1 step "
                                                   "each, one after another, %d threads" % n}}


def probe(which):
    """Launch sequence for counter runs: three copies of a known size (256 MiB
    read and written: the counter calibration), then one SSP-RK3 step of the
    headline workload (``fp64``, ``fp32-bulk``: three qline_kernel stages
    and their moment passes) or of the 2D3V electromagnetic one (``em-fp64``,
    ``em-fp32-bulk``: three qcell_kernel stages)."""
    import torch
    from tests.cases import case_2d2v, case_2d3v, case_c5, hip_solver
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    if which == "ensemble":
        # one SSP-RK3 step of the 96-member workload at the work shape the
        # autotuner picks for it on this GPU
        from bolt_amd.ensemble import Ensemble
        m = case_c5(seed=0, precision="fp32-bulk")
        ens = Ensemble([hip_solver(c) for c, _ in m], [dt for _, dt in m],
                       tuning=ENSEMBLE_PROBE_TUNING, autotune=False, tuning_cache=False)
        ens.step(1)
        torch.cuda.synchronize()
        print("probe done", flush=True)
        return
    if which.startswith("em-"):
        c = case_2d3v(seed=0, precision=which[3:])
    else:
        c = case_2d2v(seed=0, precision=which)
    nls = hip_solver(c)
    nls._advance(c.dt)
    torch.cuda.synchronize()
    print("probe done", flush=True)


def probe_algorithmic(which):
    """Algorithmic bytes of the probe's three stage launches: stage 1 reads
    u and writes out, stages 2 and 3 also read un.  The ensemble's step
    alternates stage and finalize launches; the finalize has no algorithmic
    bytes of its own (its traffic is the moment partials and tables)."""
    if which == "ensemble":
        from tests.cases import case_c5
        cells = sum(c.domain.N_q1 * c.domain.N_p1 * c.domain.N_p2 for c, _ in case_c5(seed=0))
        return [8 * cells, 0, 12 * cells, 0, 12 * cells, 0]
    cells = 64 * 64 * 24 ** 3 if which.startswith("em-") else 128 * 128 * 48 * 48
    esz = 4 if which.endswith("fp32-bulk") else 8
    return [2 * cells * esz, 3 * cells * esz, 3 * cells * esz]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    # at least 1 s of warm-up steps is run whatever this says (clocks ramp up
    # from idle)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 2D3V electromagnetic and adjoint measurements")
    ap.add_argument("--probe", choices=["fp64", "fp32-bulk", "em-fp64", "em-fp32-bulk", "ensemble"],
                    help="traffic probe for rocprofv3 --pmc: a 256 MiB device copy "
                         "(the counter calibration) and then one step of the workload")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="with --probe-of: turn the counter CSVs of a FETCH_SIZE and a "
                         "WRITE_SIZE run of that probe into profiles/traffic.json")
    ap.add_argument("--probe-of", choices=["fp64", "fp32-bulk", "em-fp64", "em-fp32-bulk",
                                           "ensemble"])
    ap.add_argument("--only", choices=["em_2d3v", "adjoint", "ensemble"],
                    help="run one of the other workloads alone (for profiling) and print "
                         "its JSON object")
    args = ap.parse_args()
    if args.traffic:
        w = args.probe_of
        if w == "ensemble":
            e = traffic_from_pmc("ensemble", args.traffic[0], args.traffic[1],
                                 "ens_chunk_kernel|ens_finalize_kernel", probe_algorithmic(w),
                                 after="ens_chunk_kernel",
                                 extra={"tuning": list(ENSEMBLE_PROBE_TUNING)})
            print(json.dumps({w: e}, indent=1), flush=True)
            return
        kernel = "qcell_kernel|qsweep_kernel" if w.startswith("em-") else "qline_kernel"
        key = "em_" + w[3:] if w.startswith("em-") else w
        e = traffic_from_pmc(key, args.traffic[0], args.traffic[1], kernel, probe_algorithmic(w))
        print(json.dumps({key: e}, indent=1), flush=True)
        return
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
    from tests.cases import case_2d2v
    copy_gbs = copy_bandwidth()
    nls64, r64 = run_gpu('fp64', args.steps, args.warmup)
    del nls64
    nls32, r32 = run_gpu('fp32-bulk', args.steps, args.warmup)
    del nls32
    # accuracy of the FP32-bulk step against the FP64 step, same inputs, same steps
    n_err = 100
    rel_l2 = fp32_bulk_error(case_2d2v, n_err)
    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "phase-space cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "ms_per_step_repeats": r64["ms_per_step_repeats"],
        "ms_per_step_without_graph": r64["ms_per_step_without_graph"],
        "timing": "each timed step is one replay of a HIP graph holding one step's launches "
                  "(nonlinear_solver.capture); the stage kernel's roofline times eager launches",
        "warmup_steps_run": r64["warmup_steps_run"],
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
                      "ms_per_step_repeats": r32["ms_per_step_repeats"],
                      "ms_per_step_without_graph": r32["ms_per_step_without_graph"],
                      "dtype": "f32 bulk, f64 moments/temperature",
                      "roofline": dict(r32["roofline"], copy_gbs=copy_gbs,
                                       frac_of_copy=r32["roofline"]["achieved"] / copy_gbs),
                      "rel_l2_vs_fp64_after_%d_steps" % n_err: rel_l2,
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
