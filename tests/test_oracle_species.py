"""Pins of the oracle's multi-species path (CPU): species are independent
when the fields are off (each equals its own one-species run), and the
fields see the species sum of rho (two equal species of opposite charge
leave E at zero)."""
import pytest
import torch

from oracle import bolt_oracle as bo
from tests.cases import case_species, oracle_solver, single_species_of


def _run(c, steps):
    o = oracle_solver(c)
    for _ in range(steps):
        o.strang_timestep(c.dt)
    return o


def test_species_are_independent_without_fields():
    c = case_species(stepper="SSP-RK3")
    o = _run(c, 4)
    for k in range(2):
        ok = _run(single_species_of(c, k), 4)
        a = o.interior()[:, k:k + 1]
        b = ok.interior()
        assert float((a - b).abs().max()) <= 1e-15 * float(b.abs().max()), k


def test_opposite_charges_leave_the_field_at_zero():
    """Two species with the same mass, temperature, drift, grid and density
    but charges +1 and -1: rho sums to zero, so E stays zero and each species
    streams as with the fields off."""
    kw = dict(masses=(1.0, 1.0), charges=(1.0, -1.0), temps=(1.0, 1.0), drifts=(0.3, 0.3),
              p_max=(6.0, 6.0), taus=(0.05, 0.05), N_q2=1, N_p2=1, p_dim=1, N_p1=12,
              recon_p="minmod", amp=5e-2)
    c = case_species(fields=True, fields_type="electrostatic", **kw)
    # the same density phase for both species
    f_all = c.initialize.initialize_f
    c.initialize.initialize_f = lambda q1, q2, p1, p2, p3, params: torch.cat(
        [f_all(q1, q2, p1, p2, p3, params)[:, :1]] * 2, dim=1)
    o = _run(c, 5)
    E = o.fields_solver.cell_centered_EM_fields
    assert float(E.abs().max()) <= 1e-14
    off = case_species(fields=False, **kw)
    off.initialize.initialize_f = c.initialize.initialize_f
    o2 = _run(off, 5)
    assert float((o.interior() - o2.interior()).abs().max()) <= 1e-15 * float(o2.interior().abs().max())


def test_species_pins_catch_planted_mistakes(monkeypatch):
    from tests.test_oracle_pins import _mutant
    # moments summed over species before BGK (a species-mixing mistake)
    monkeypatch.setattr(bo, "integral_over_v", _mutant(
        bo.integral_over_v, "array.sum(dim=0, keepdim=True)",
        "array.sum(dim=0, keepdim=True).sum(dim=1, keepdim=True)"))
    with pytest.raises(AssertionError):
        test_species_are_independent_without_fields()
    monkeypatch.undo()
    # rho taken from the first species only
    monkeypatch.setattr(bo.OracleFields, "compute_electrostatic_fields", _mutant(
        bo.OracleFields.compute_electrostatic_fields, "rho = rho.sum(dim=1, keepdim=True)",
        "rho = rho[:, :1]"))
    with pytest.raises(AssertionError):
        test_opposite_charges_leave_the_field_at_zero()
