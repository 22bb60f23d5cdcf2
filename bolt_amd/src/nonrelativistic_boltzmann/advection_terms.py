"""Advection terms, bolt/src/nonrelativistic_boltzmann/advection_terms.py:28-188.

Arguments follow the reference: q-arrays shaped (1, 1, N_q1, N_q2) and
p-arrays shaped (N_p, N_s, 1, 1) (torch tensors, broadcasting like
af.broadcast).
"""


def A_q(t, q1, q2, v1, v2, v3, params):
    """advection_terms.py:28-60."""
    return (v1, v2)


def C_q(t, q1, q2, v1, v2, v3, params):
    """advection_terms.py:64-97: C_q1 = v1, C_q2 = v2."""
    return (v1, v2)


def A_p(t, q1, q2, v1, v2, v3, fields_solver, params):
    """advection_terms.py:99-143."""
    return C_p(t, q1, q2, v1, v2, v3, fields_solver, params)


def C_p(t, q1, q2, v1, v2, v3, fields_solver, params):
    """advection_terms.py:147-188: Lorentz force with cell-centred fields from
    fields_solver.get_fields() (which alternates between the n and n+1/2 time
    levels on every call)."""
    e = params.charge
    m = params.mass
    E1, E2, E3, B1, B2, B3 = fields_solver.get_fields()
    C_p1 = (e / m) * (E1 + v2 * B3 - v3 * B2)
    C_p2 = (e / m) * (E2 + v3 * B1 - v1 * B3)
    C_p3 = (e / m) * (E3 + v1 * B2 - v2 * B1)
    return (C_p1, C_p2, C_p3)
