"""Committed oracle fixtures (tests/golden/oracle_fixtures.npz): the oracle
must still produce them (CPU), and the HIP path must match them (GPU)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import HERE, golden_cases, run_oracle

FIX = np.load(os.path.join(HERE, "oracle_fixtures.npz"))
NAMES = sorted(golden_cases())


def _rel_linf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    f0, f1 = run_oracle(name)
    assert _rel_linf(f0, FIX[name + "__in"]) <= 1e-15
    assert _rel_linf(f1, FIX[name + "__out"]) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_fixture(gpu, name):
    from tests.cases import hip_solver
    make, steps = golden_cases()[name]
    c = make()
    nls = hip_solver(c)
    g1, g2 = nls.N_g1, nls.N_g2

    def interior():
        return nls.f[:, :, g1:g1 + nls.N_q1, g2:g2 + nls.N_q2].detach().cpu().double().numpy()
    assert _rel_linf(interior(), FIX[name + "__in"]) <= 1e-15
    for _ in range(steps):
        nls.strang_timestep(c.dt)
    err = _rel_linf(interior(), FIX[name + "__out"])
    print(f"{name}: rel Linf {err:.3e}", flush=True)
    assert err <= (1e-11 if steps > 1 else 1e-12)
