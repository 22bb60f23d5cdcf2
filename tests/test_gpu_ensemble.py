"""The ensemble step on the GPU: members of uneven sizes advanced together
against the oracle (one step from the same state; the bar is 1e-12 or 4x the
oracle's own change under 1-ulp input noise, as for the single-problem
kernels) and against the single-problem solver, FP32-bulk against FP64,
invariance of the results under the work-list shape, the autotuner, and the
divergence flag."""
import numpy as np
import pytest
import torch

from tests.cases import case_c5, hip_solver, maxwellian_case, oracle_solver

pytestmark = pytest.mark.gpu


def _members(precision, sizes, stepper="SSP-RK3", seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for n1, npp in sizes:
        c = maxwellian_case(N_q1=n1, N_q2=1, N_p1=npp, N_p2=npp, p_dim=2, p_max=10.0,
                            modes=((int(rng.integers(1, 4)), 0),),
                            amps=(float(rng.uniform(1e-3, 1e-2)),),
                            phases=(float(rng.uniform(0, 6.28)),), recon_q="weno5",
                            recon_p="weno5", riemann="lax-friedrichs", stepper=stepper,
                            tau=float(10 ** rng.uniform(-2, 0)), precision=precision)
        c.dt = 0.4 / n1 / (10.0 - 10.0 / npp)
        out.append(c)
    return out


SIZES = [(32, 8), (40, 6), (64, 10), (128, 6), (256, 4), (48, 12), (37, 6), (16, 8)]


def _rel_linf(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("stepper", ["SSP-RK3", "RK2"])
def test_ensemble_one_step_vs_oracle(gpu, stepper):
    from bolt_amd.ensemble import Ensemble
    cases = _members("fp64", SIZES, stepper)
    ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    ens.step(1)
    torch.cuda.synchronize()
    from tests.test_gpu_parity import _oracle_self_sensitivity
    for i, c in enumerate(cases):
        sens = _oracle_self_sensitivity(c)
        o = oracle_solver(c)
        o.strang_timestep(c.dt)
        ref = o.interior()[:, 0, :, 0]  # (P, N_q1)
        err = _rel_linf(ens.state(i).cpu(), ref)
        print(f"member {i} {SIZES[i]} {stepper}: rel Linf {err:.3e} (oracle sensitivity "
              f"{sens:.3e})", flush=True)
        assert err <= max(1e-12, 4 * sens)


def test_ensemble_matches_single_solver_runs(gpu):
    from bolt_amd.ensemble import Ensemble
    cases = _members("fp64", SIZES)
    solo = [hip_solver(c) for c in cases]
    ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    ens.check_divergence()  # finite initial state: the flag must be clear
    ens.step(10)
    for s, c in zip(solo, cases):
        for _ in range(10):
            s.strang_timestep(c.dt)
    ens.check_divergence()
    worst = 0.0
    for i, s in enumerate(solo):
        worst = max(worst, _rel_linf(ens.state(i), ens._lines(s)))
        m = ens.moments(i)[0]
        mref = s._mom[0]
        assert float((m - mref).abs().max() / mref.abs().max()) <= 1e-10
    print(f"ensemble vs single-problem solver after 10 steps: rel Linf {worst:.3e}", flush=True)
    assert worst <= 1e-10


@pytest.mark.parametrize("recon", ["minmod", "ppm", "piecewise-constant"])
def test_ensemble_other_reconstructions_match_single_solver(gpu, recon):
    """The chunk kernel's minmod, PPM and piecewise-constant forms (odd line
    lengths take the one-cell path) against the single-problem solver."""
    from bolt_amd.ensemble import Ensemble
    rng = np.random.default_rng(4)
    cases = []
    for n1, npp in [(32, 8), (37, 6), (128, 6), (16, 8)]:
        c = maxwellian_case(N_q1=n1, N_q2=1, N_p1=npp, N_p2=npp, p_dim=2, p_max=10.0,
                            modes=((1, 0),), amps=(float(rng.uniform(1e-3, 1e-2)),),
                            phases=(float(rng.uniform(0, 6.28)),), recon_q=recon,
                            recon_p=recon, riemann="upwind-flux", stepper="SSP-RK3",
                            tau=0.05, precision="fp64")
        c.dt = 0.4 / n1 / 10.0
        cases.append(c)
    solo = [hip_solver(c) for c in cases]
    ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    ens.step(5)
    for s, c in zip(solo, cases):
        for _ in range(5):
            s.strang_timestep(c.dt)
    worst = max(_rel_linf(ens.state(i), ens._lines(s)) for i, s in enumerate(solo))
    print(f"{recon}: ensemble vs single-problem solver after 5 steps: rel Linf {worst:.3e}",
          flush=True)
    assert worst <= 1e-10


def test_ensemble_fp32_bulk_vs_fp64(gpu):
    from bolt_amd.ensemble import Ensemble
    c64 = _members("fp64", SIZES, seed=3)
    c32 = _members("fp32-bulk", SIZES, seed=3)
    e64 = Ensemble([hip_solver(c) for c in c64], [c.dt for c in c64], autotune=False)
    e32 = Ensemble([hip_solver(c) for c in c32], [c.dt for c in c32], autotune=False)
    e64.step(20)
    e32.step(20)
    for i in range(len(SIZES)):
        a, b = e32.state(i).double(), e64.state(i)
        l2 = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        assert l2 <= 1e-5, (i, l2)


def test_batch_size_does_not_change_results(gpu):
    """The cells a workgroup stages at once (one line per batch up to whole
    chunks, with a partial last batch) leave the results bit-identical; a
    different chunk size regroups only the FP64 moment sums."""
    from bolt_amd.ensemble import Ensemble
    cases = _members("fp64", SIZES, seed=5)
    outs = []
    for tuning in [(16384, 256), (16384, 2048), (16384, 1024), (16384, 1)]:
        ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], tuning=tuning)
        ens.step(3)
        outs.append(ens.f.clone())
    for o in outs[1:]:
        assert torch.equal(outs[0], o)
    ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], tuning=(512, 2048))
    ens.step(3)
    d = _rel_linf(ens.f, outs[0])
    print(f"chunk 512 vs 16384 cells after 3 steps: rel Linf {d:.3e}", flush=True)
    assert d <= 1e-10


def test_autotune_keeps_state_and_picks_a_candidate(gpu):
    from bolt_amd.ensemble import Ensemble
    members = case_c5(seed=1, n=12)
    ens = Ensemble([hip_solver(c) for c, _ in members], [dt for _, dt in members],
                   autotune=False)
    f0 = ens.f.clone()
    cands = [(8192, 1024), (16384, 2048)]
    best = ens.autotune(candidates=cands, steps=1)
    assert best in cands
    ens._set_tuning(best)
    assert torch.equal(ens.f, f0)
    ens.step(2)
    ens.check_divergence()


def test_divergence_flag(gpu):
    from bolt_amd.ensemble import Ensemble
    cases = _members("fp64", SIZES[:3], seed=7)
    ens = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    ens.state(1)[2, 3] = float("nan")
    ens.step(1)
    with pytest.raises(SystemExit):
        ens.check_divergence()


def test_graph_replays_match_eager_steps(gpu):
    """Steps captured in a HIP graph and replayed three times, with eager work
    and synchronisations between the replays, equal eager stepping bit for
    bit after every replay."""
    from bolt_amd.ensemble import Ensemble
    cases = _members("fp64", SIZES, seed=9)
    eager = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    graph = Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], autotune=False)
    graph.capture(2)
    for _ in range(3):
        eager.step(2)
        graph.replay()
        scratch = torch.ones(1024, device=gpu).cumsum(0)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(eager.f, graph.f)
    graph.check_divergence()


def test_tuning_cache_is_read_instead_of_tuning_again(gpu, tmp_path, monkeypatch):
    """The autotuned shape is stored in the JSON cache under the ensemble's
    signature and the GPU model; a later ensemble with the same signature
    (its in-process cache cleared, as in a new process) takes it from the
    file and does not tune."""
    import os
    from bolt_amd import ensemble as E
    cases = _members("fp32-bulk", SIZES[:4])
    path = str(tmp_path / "tuning.json")
    E._TUNED.clear()
    a = E.Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], tuning_cache=path)
    assert os.path.exists(path)
    assert E.load_tuning(path, a._cache_key()) == a.tuning
    E._TUNED.clear()

    def no_tuning(self, *args, **kw):
        raise AssertionError("tuned again despite the cache")
    monkeypatch.setattr(E.Ensemble, "autotune", no_tuning)
    b = E.Ensemble([hip_solver(c) for c in cases], [c.dt for c in cases], tuning_cache=path)
    assert b.tuning == a.tuning
    E._TUNED.clear()
