"""GPU parity of the non-periodic q boundaries (mirror, dirichlet,
mirror+dirichlet; boundaries.py:147-577) against the FP64 oracle, through the
C-ABI (boltx_fill_ghosts + the stage kernels reading the ghost rows)."""
import numpy as np
import pytest
import torch

from tests.cases import case_bc, hip_solver, oracle_solver
from tests.test_gpu_parity import _cmp, _per_step

pytestmark = pytest.mark.gpu

CASES = {
    "mirror4": dict(left="mirror", right="mirror", bottom="mirror", top="mirror"),
    "dirichlet4": dict(left="dirichlet", right="dirichlet", bottom="dirichlet", top="dirichlet"),
    "mirror+dirichlet_q1": dict(left="mirror+dirichlet", right="mirror+dirichlet"),
    "mixed": dict(left="mirror", right="dirichlet", bottom="mirror+dirichlet",
                  top="mirror"),
    "q2_only": dict(left="periodic", right="periodic", bottom="dirichlet", top="mirror"),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("rq", ["piecewise-constant", "minmod", "ppm", "weno5"])
def test_bcs_per_step(gpu, name, rq):
    stepper = "SSP-RK3" if rq in ("ppm", "weno5") else "RK2"
    riemann = "lax-friedrichs" if rq == "minmod" else "upwind-flux"
    c = case_bc(recon_q=rq, stepper=stepper, riemann=riemann, source=(rq != "ppm"),
                **CASES[name])
    print(f"case {name} {rq} {stepper}", flush=True)
    _per_step(c, 3, 1e-12)


def test_bcs_full_run_and_moments(gpu):
    c = case_bc(recon_q="weno5", stepper="SSP-RK3", source=True, **CASES["mixed"])
    o, nls = oracle_solver(c), hip_solver(c)
    for _ in range(12):
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
    linf, l2 = _cmp(nls, o)
    print(f"mixed BCs 12 steps: rel Linf {linf:.3e} rel L2 {l2:.3e}", flush=True)
    assert linf <= 1e-11
    g1, g2 = nls.N_g1, nls.N_g2
    for name in ("density", "mom_v1_bulk", "mom_v2_bulk", "energy"):
        a = nls.compute_moments(name).cpu().numpy()[..., g1:g1 + nls.N_q1, g2:g2 + nls.N_q2]
        b = o.compute_moments(name).numpy()[..., g1:g1 + nls.N_q1, g2:g2 + nls.N_q2]
        assert float(np.abs(a - b).max() / np.abs(b).max()) <= 1e-11, name


def test_beam_test_1d_configuration(gpu):
    """The reference's beam_test/1D set-up (example_problems/
    nonrelativistic_boltzmann/beam_test/1D: 32 q1 cells with dirichlet
    walls, 3 periodic q2 cells, 4 p1 cells on [-4, 4], minmod, upwind, no
    source) with inflowing Maxwellians on both sides."""
    c = case_bc(left="dirichlet", right="dirichlet", bottom="periodic", top="periodic",
                N_q1=32, N_q2=3, N_p1=4, p_dim=1, p_max=4.0, recon_q="minmod",
                stepper="RK2", source=False, dt=1e-3, drift=(0.5, 0.0))
    _per_step(c, 5, 1e-12)
    o, nls = oracle_solver(c), hip_solver(c)
    for _ in range(50):
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
    linf, _ = _cmp(nls, o)
    print(f"beam 1D 50 steps: rel Linf {linf:.3e}", flush=True)
    assert linf <= 1e-11


def test_ghost_zones_match_the_oracle_refresh(gpu):
    """nls.f shows the ghost zones the boundary conditions fill; they equal
    the oracle's after communicate_f + apply_bcs_f (corners excluded: no
    stencil reads them)."""
    c = case_bc(recon_q="weno5", **CASES["mixed"])
    o, nls = oracle_solver(c), hip_solver(c)
    for _ in range(2):
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
    o._communicate_f()
    a = nls.f.cpu().numpy()
    b = o.f.numpy()
    g = nls.N_ghost
    n1, n2 = nls.N_q1, nls.N_q2
    mask = np.ones(a.shape[2:], dtype=bool)
    mask[:g, :g] = mask[:g, -g:] = mask[-g:, :g] = mask[-g:, -g:] = False
    d = np.abs(a - b)[..., mask].max() / np.abs(b).max()
    assert d <= 1e-12, d
    assert np.abs(b[:, :, :g, g:g + n2]).max() > 0  # the ghost zone is not empty


def test_fp32_bulk_with_bcs(gpu):
    c64 = case_bc(recon_q="weno5", stepper="SSP-RK3", source=True, **CASES["mixed"])
    c32 = case_bc(recon_q="weno5", stepper="SSP-RK3", source=True, precision="fp32-bulk",
                  **CASES["mixed"])
    a, b = hip_solver(c64), hip_solver(c32)
    for _ in range(10):
        a.strang_timestep(c64.dt)
        b.strang_timestep(c32.dt)
    x, y = a.f.double(), b.f.double()
    rel = float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(x))
    print(f"fp32-bulk vs fp64 with BCs: rel L2 {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_mirror_box_conserves_mass_at_size(gpu):
    """A 64 x 64 q box with mirror walls, 16 x 16 p, PPM, 20 steps: the total
    of f is conserved to rounding (the oracle pin of the same property runs
    on the CPU at a small size)."""
    c = case_bc(left="mirror", right="mirror", bottom="mirror", top="mirror", N_q1=64,
                N_q2=64, N_p1=16, N_p2=16, recon_q="ppm", stepper="SSP-RK3", source=False,
                dt=1e-3, drift=(1.0, -0.7))
    nls = hip_solver(c)
    g = nls.N_ghost

    def mass():
        return float(nls.f[:, :, g:g + 64, g:g + 64].double().sum())

    m0 = mass()
    for _ in range(20):
        nls.strang_timestep(c.dt)
    m1 = mass()
    print(f"mirror box mass drift {abs(m1 - m0) / m0:.3e}", flush=True)
    assert abs(m1 - m0) <= 1e-13 * m0


@pytest.mark.parametrize("name", ["mirror4", "dirichlet4", "mixed", "q2_only"])
@pytest.mark.parametrize("rq", ["minmod", "weno5"])
def test_bcs_electrostatic_per_step(gpu, name, rq):
    """Non-periodic q boundaries with electrostatic fields (FFT Poisson from
    the interior density, fft.py:33, p-advection by the cell-centred fields):
    the field ghost values apply_bcs_fields sets never reach an interior
    update, so the interior follows the oracle per step."""
    stepper = "SSP-RK3" if rq == "weno5" else "RK2"
    c = case_bc(recon_q=rq, stepper=stepper, source=True, electrostatic=True, **CASES[name])
    print(f"case {name} {rq} {stepper} electrostatic", flush=True)
    _per_step(c, 3, 1e-12)


def test_bcs_electrostatic_slab_and_generic_kernels_agree(gpu):
    """The same ghosted electrostatic stage through the slab-ordered kernel
    (with the workspace) and the generic one (without): equal to rounding."""
    c = case_bc(recon_q="weno5", stepper="SSP-RK3", source=True, electrostatic=True,
                **CASES["mixed"])
    a, b = hip_solver(c), hip_solver(c)
    b._spc[0].ws = None
    for _ in range(3):
        a.strang_timestep(c.dt)
        b.strang_timestep(c.dt)
    d = float((a._f - b._f).abs().max() / a._f.abs().max())
    assert d <= 1e-13, d
