"""Writes the regression fixtures in this folder from the FP64 oracle.

The reference ships no golden vectors, so these are oracle outputs on small
seeded problems: the 1D1V minmod/RK2 case after 1 and 100 steps, one step of a
small 2D2V problem per reconstruction x stepper, one and five steps of a small
2D3V electromagnetic problem under SSP-RK3 (fields advanced once per step) and
three under RK2 (the reference's field coupling), ten steps of a 1D1V electrostatic (FFT Poisson)
problem, five steps of a 2D2V box with mirror, dirichlet and mirror+dirichlet
walls, and five steps of a two-species 2D2V problem.  They pin the oracle against accidental change and
give the GPU tests fixed expected outputs.  Run from the repository root:

    python -m tests.golden.make_golden
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_cases():
    """name -> (case factory, steps)."""
    from tests.cases import case_2d2v, case_2d3v, case_bc, case_c1, case_es, case_species
    out = {"c1_1step": (lambda: case_c1(seed=0), 1), "c1_100steps": (lambda: case_c1(seed=0), 100)}
    for rq in ("piecewise-constant", "minmod", "ppm", "weno5"):
        for st in ("RK2", "SSP-RK3"):
            out[f"2d2v_{rq}_{st}"] = (
                lambda rq=rq, st=st: _dt(case_2d2v(seed=1, N_q=8, N_p=6, recon_q=rq,
                                                   recon_p=rq, stepper=st), 1e-3), 1)
    out["2d3v_em"] = (lambda: _dt(case_2d3v(seed=2, N_q=6, N_p=4), 1e-3), 1)
    out["2d3v_em_rk2_3steps"] = (lambda: _dt(case_2d3v(seed=4, N_q=6, N_p=4, stepper="RK2"),
                                             1e-3), 3)
    out["2d3v_em_5steps"] = (lambda: _dt(case_2d3v(seed=5, N_q=6, N_p=4), 1e-3), 5)
    out["1d1v_electrostatic_10steps"] = (lambda: case_es(seed=3), 10)
    out["2d2v_walls_5steps"] = (lambda: case_bc(
        left="mirror", right="dirichlet", bottom="mirror+dirichlet", top="mirror",
        recon_q="weno5", stepper="SSP-RK3", source=True), 5)
    out["2d2v_two_species_5steps"] = (lambda: case_species(stepper="SSP-RK3"), 5)
    return out


def _dt(case, dt):
    case.dt = dt
    return case


def run_oracle(name):
    from tests.cases import oracle_solver
    make, steps = golden_cases()[name]
    c = make()
    o = oracle_solver(c)
    f0 = o.interior().clone().numpy()
    for _ in range(steps):
        o.strang_timestep(c.dt)
    return f0, o.interior().numpy()


def main():
    arrays = {}
    for name in golden_cases():
        f0, f1 = run_oracle(name)
        arrays[name + "__in"] = f0
        arrays[name + "__out"] = f1
    np.savez_compressed(os.path.join(HERE, "oracle_fixtures.npz"), **arrays)


if __name__ == "__main__":
    main()
