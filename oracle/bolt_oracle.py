"""FP64 restatement of Bolt's nonlinear FVM time step (TEST INFRASTRUCTURE ONLY).

Arrays keep the reference's logical axis order ``(P, S, Q1g, Q2g)`` (ArrayFire
column-major, p1 fastest inside P) as torch float64 CPU tensors, so every
``af.shift`` / ``af.moddims`` / ``af.sum`` of the reference has a one-line
counterpart here.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this module.

What pins each function (the tests live in ``tests/test_oracle_*.py``):

=======================  =====================================================
function                 pinned by
=======================  =====================================================
af_shift / moddims       ArrayFire's documented semantics, checked on index
                         arrays (out[i] = in[i-k], column-major reshape) and
                         the reference's layout test ``test_convert.py:51-84``
calculate_q / _p         closed forms (cell centres ``start+(i+1/2)d``) and the
                         reference's ``test_calculate_p.py`` / ``_q.py`` values
reconstruct_minmod       closed form minmod(2b, c/2, 2f); exact on linear data,
                         zero slope at extrema, brute force on 3-cell stencils
reconstruct_ppm          exact on linear data, constants preserved, mirror
                         symmetry, no new extrema on monotone data
reconstruct_weno5        exact on quadratics, ideal weights reproduce the 5-pt
                         Lagrange interpolant (exact on quartics), the
                         expanded beta equals the Jiang-Shu sum of squares
riemann (upwind / LLF)   LLF == upwind to rounding for a linear flux; C=0 ->
                         zero flux; select picks the right state at C == 0
df_dt_fvm                periodic conservation (sum of df/dt = 0), constant
                         state is stationary, DFT eigenvalue of a sine mode
ghosted_fill /           mirror walls conserve mass exactly and keep a
apply_bcs_f              mirror-symmetric state symmetric; piecewise-constant
                         + upwind with dirichlet inflow equals the first-order
                         upwind scheme with inflow values, written from its
                         definition; a uniform state equal to the inflow
                         value is stationary
compute_moments          the reference KAT ``test_compute_moments.py:76-144``
                         (drifting Maxwellian, 1e-13)
BGK / f0                 C[f_MB] ~ 0, conservation of n, nu, E, relaxation
                         test of ``collision_operator_test/main.py:62-97``
RK2 / SSP-RK3            convergence order on df/dt = f
                         (``test_time_integrators.py:26-70``)
FDTD                     divB preserved to rounding, plane-wave mode order
                         (``test_fdtd_explicit.py``), discrete energy
=======================  =====================================================
"""
import math
import types

import torch

F64 = torch.float64


# ---------------------------------------------------------------------------
# ArrayFire primitives
# ---------------------------------------------------------------------------
def af_shift(x, *shifts):
    """``af.shift(x, k0, k1, k2, k3)``: circular, ``out[i] = x[i - k]``."""
    dims = [d for d, k in enumerate(shifts) if k != 0]
    ks = [int(k) for k in shifts if k != 0]
    if not dims:
        return x
    return torch.roll(x, shifts=ks, dims=dims)


def moddims(x, *dims):
    """``af.moddims``: reinterpret the column-major buffer with new dims."""
    nd = x.dim()
    flat = x.permute(*range(nd - 1, -1, -1)).reshape(-1)
    return flat.reshape(*dims[::-1]).permute(*range(len(dims) - 1, -1, -1))


def as4(x):
    """Scalar or tensor -> float64 tensor (broadcast like af.broadcast)."""
    if isinstance(x, torch.Tensor):
        return x.to(F64)
    return torch.tensor(float(x), dtype=F64)


# ---------------------------------------------------------------------------
# Grids: bolt/lib/utils/calculate_q.py:7-41, calculate_p.py:8-57
# ---------------------------------------------------------------------------
def calculate_q(q1_start, q2_start, N_q1, N_q2, N_g1, N_g2, dq1, dq2):
    """calculate_q.py:7-41.  Returns (left_bot, center_bot, left_center, center)
    pairs of (1,1,Q1g,Q2g) arrays; ``left_center``'s q2 entry is q2_center_bot
    exactly as in calculate_q.py:37."""
    i_q1 = torch.arange(-N_g1, N_q1 + N_g1, dtype=F64)
    i_q2 = torch.arange(-N_g2, N_q2 + N_g2, dtype=F64)
    q1_left_bot = q1_start + i_q1 * dq1
    q2_left_bot = q2_start + i_q2 * dq2
    Q1 = q1_left_bot.reshape(1, 1, -1, 1).expand(1, 1, len(i_q1), len(i_q2))
    Q2 = q2_left_bot.reshape(1, 1, 1, -1).expand(1, 1, len(i_q1), len(i_q2))
    q1_left_bot, q2_left_bot = Q1.clone(), Q2.clone()
    q1_center_bot = q1_left_bot + 0.5 * dq1
    q2_center_bot = q2_left_bot
    q1_left_center = q1_left_bot
    q2_left_center = q2_left_bot + 0.5 * dq2  # computed but not returned (:37)
    q1_center = q1_left_bot + 0.5 * dq1
    q2_center = q2_left_bot + 0.5 * dq2
    del q2_left_center
    return ([q1_left_bot, q2_left_bot], [q1_center_bot, q2_center_bot],
            [q1_left_center, q2_center_bot], [q1_center, q2_center])


def _p_mesh(starts, N, d, offset):
    """numpy.meshgrid(p2, p1, p3) flattened column-major (calculate_p.py:17-26):
    p-index = i1 + N1*(i2 + N2*i3), p1 fastest."""
    S = len(starts[0])
    P = N[0] * N[1] * N[2]
    out = [torch.zeros(P, S, dtype=F64) for _ in range(3)]
    for s in range(S):
        axes = [starts[k][s] + (offset + torch.arange(N[k], dtype=F64)) * d[k][s]
                for k in range(3)]
        g3, g2, g1 = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        out[0][:, s] = g1.reshape(-1)
        out[1][:, s] = g2.reshape(-1)
        out[2][:, s] = g3.reshape(-1)
    return [o.reshape(P, S, 1, 1) for o in out]


def calculate_p_center(p1_start, p2_start, p3_start, N_p1, N_p2, N_p3, dp1, dp2, dp3):
    """calculate_p.py:33-57: ``p = start + (0.5 + i) * dp``."""
    return _p_mesh([p1_start, p2_start, p3_start], [N_p1, N_p2, N_p3],
                   [dp1, dp2, dp3], 0.5)


def calculate_p_corner(p1_start, p2_start, p3_start, N_p1, N_p2, N_p3, dp1, dp2, dp3):
    """calculate_p.py:8-31: ``p = start + i * dp``."""
    return _p_mesh([p1_start, p2_start, p3_start], [N_p1, N_p2, N_p3],
                   [dp1, dp2, dp3], 0.0)


# ---------------------------------------------------------------------------
# Reconstruction: finite_volume/reconstruction/{minmod,ppm,weno5}.py
# ---------------------------------------------------------------------------
def _af_sign(x):
    """``af.sign``: 1 for x < 0, 0 otherwise (minmod.py:8)."""
    return (x < 0).to(F64)


def minmod(x, y, z):
    """minmod.py:4-16."""
    min_of_all = torch.minimum(torch.minimum(x.abs(), y.abs()), z.abs())
    signx = 1 - 2 * _af_sign(x)
    signy = 1 - 2 * _af_sign(y)
    signz = 1 - 2 * _af_sign(z)
    return 0.25 * (signx + signy).abs() * (signx + signz) * min_of_all


def _shift_axis(x, axis, k):
    s = [0, 0, 0, 0]
    s[axis] = k
    return af_shift(x, *s)


def slope_minmod(x, axis):
    """minmod.py:18-62 (MC limiter, theta = 2)."""
    f_i_plus_one = _shift_axis(x, axis, -1)
    f_i_minus_one = _shift_axis(x, axis, 1)
    forward_diff = f_i_plus_one - x
    backward_diff = x - f_i_minus_one
    central_diff = backward_diff + forward_diff
    theta = 2
    return minmod(theta * backward_diff, 0.5 * central_diff, theta * forward_diff)


def reconstruct_minmod(x, axis):
    """minmod.py:64-71."""
    slope = slope_minmod(x, axis)
    return x - 0.5 * slope, x + 0.5 * slope


def _five(x, axis):
    """y0..y4 = x[i-2..i+2] along ``axis`` (ppm.py:51-55, weno5.py:51-55)."""
    return (_shift_axis(x, axis, 2), _shift_axis(x, axis, 1), x,
            _shift_axis(x, axis, -1), _shift_axis(x, axis, -2))


def ppm_faces(y0, y1, y2, y3, y4):
    """ppm.py:57-98 on explicit stencil values; returns (left, right)."""
    d0 = 2 * (y1 - y0)
    d1 = 2 * (y2 - y1)
    d2 = 2 * (y3 - y2)
    d3 = 2 * (y4 - y3)
    D1 = 0.5 * (y2 - y0)
    D2 = 0.5 * (y3 - y1)
    D3 = 0.5 * (y4 - y2)

    def dq(dl, dr, D):
        cond = (dr * dl <= 0).to(F64)
        sign = (D > 0).to(F64) * 2 - 1
        return (1 - cond) * sign * torch.minimum(D.abs(), torch.minimum(dl.abs(), dr.abs()))

    DQ1 = dq(d0, d1, D1)
    DQ2 = dq(d1, d2, D2)
    DQ3 = dq(d2, d3, D3)
    left = 0.5 * (y2 + y1) - (DQ2 - DQ1) / 6
    right = 0.5 * (y3 + y2) - (DQ3 - DQ2) / 6
    corr1 = ((right - y2) * (y2 - left) <= 0).to(F64)
    qd = right - left
    qe = 6 * (y2 - 0.5 * (right + left))
    corr2 = (qd * (qd - qe) < 0).to(F64)
    corr3 = (qd * (qd + qe) < 0).to(F64)
    left = left * (1 - corr1) + corr1 * y2
    right = right * (1 - corr1) + corr1 * y2
    left = left * (1 - corr2) + corr2 * (3 * y2 - 2 * right)
    right = (right * corr2 + (1 - corr2) * right * (1 - corr3)
             + (1 - corr2) * corr3 * (3 * y2 - 2 * left))
    return left, right


def reconstruct_ppm(x, axis):
    """ppm.py:4-100."""
    return ppm_faces(*_five(x, axis))


WENO_EPS = 1e-17


def weno5_beta(a, b, c):
    """weno5.py:58-71: expanded smoothness indicator plus the eps floor."""
    return (((4 / 3) * a * a - (19 / 3) * a * b + (25 / 3) * b * b + (11 / 3) * a * c
             - (31 / 3) * b * c + (10 / 3) * c * c)
            + WENO_EPS * (1.0 + a.abs() + b.abs() + c.abs()))


def weno5_faces(y0, y1, y2, y3, y4):
    """weno5.py:57-98 (point-value WENO-JS as written); returns (left, right)."""
    beta1 = weno5_beta(y0, y1, y2)
    beta2 = weno5_beta(y1, y2, y3)
    beta3 = weno5_beta(y2, y3, y4)
    w1r = 1 / (16 * beta1 * beta1)
    w2r = 5 / (8 * beta2 * beta2)
    w3r = 5 / (16 * beta3 * beta3)
    w1l = 5 / (16 * beta1 * beta1)
    w2l = 5 / (8 * beta2 * beta2)
    w3l = 1 / (16 * beta3 * beta3)
    denl = w1l + w2l + w3l
    denr = w1r + w2r + w3r
    u1r = 0.375 * y0 - 1.25 * y1 + 1.875 * y2
    u2r = -0.125 * y1 + 0.75 * y2 + 0.375 * y3
    u3r = 0.375 * y2 + 0.75 * y3 - 0.125 * y4
    u1l = -0.125 * y0 + 0.75 * y1 + 0.375 * y2
    u2l = 0.375 * y1 + 0.75 * y2 - 0.125 * y3
    u3l = 1.875 * y2 - 1.25 * y3 + 0.375 * y4
    left = (w1l * u1l + w2l * u2l + w3l * u3l) / denl
    right = (w1r * u1r + w2r * u2r + w3r * u3r) / denr
    return left, right


def reconstruct_weno5(x, axis):
    """weno5.py:4-98."""
    return weno5_faces(*_five(x, axis))


def reconstruct(x, axis, method):
    """reconstruct.py:7-51.  The ``SystemExit`` guard for WENO5
    (reconstruct.py:39-41) is bypassed: WENO5 is used as written in weno5.py."""
    if method == "piecewise-constant":
        return x, x
    if method == "minmod":
        return reconstruct_minmod(x, axis)
    if method == "ppm":
        return reconstruct_ppm(x, axis)
    if method == "weno5":
        return reconstruct_weno5(x, axis)
    raise NotImplementedError("Reconstruction method invalid/not-implemented")


# ---------------------------------------------------------------------------
# Riemann solvers: finite_volume/riemann.py:8-58 (+ LLF, not in the reference)
# ---------------------------------------------------------------------------
def riemann_upwind_state(left_state, right_state, velocity):
    """riemann.py:46-49: ``select(velocity > 0, left_state, right_state)``."""
    velocity = torch.broadcast_to(velocity, left_state.shape)
    return torch.where(velocity > 0, left_state, right_state)


def face_flux(method, left_state, right_state, velocity):
    """Upwind: ``C * select(C > 0, fL, fR)`` (riemann.py:46-49 +
    df_dt_fvm.py:82-83).  Local Lax-Friedrichs (added; the reference has no
    LLF although its docs name one): ``0.5*(C*(fL+fR) - |C|*(fR-fL))``."""
    if method in ("upwind-flux", None):
        return velocity * riemann_upwind_state(left_state, right_state, velocity)
    if method == "lax-friedrichs":
        return 0.5 * (velocity * (left_state + right_state)
                      - velocity.abs() * (right_state - left_state))
    raise NotImplementedError("Riemann solver invalid/not-implemented")


# ---------------------------------------------------------------------------
# Built-in physics: bolt/src/nonrelativistic_boltzmann/*.py
# ---------------------------------------------------------------------------
def C_q(t, q1, q2, v1, v2, v3, params):
    """advection_terms.py:64-97."""
    return v1, v2


def A_q(t, q1, q2, v1, v2, v3, params):
    """advection_terms.py:28-60."""
    return v1, v2


def C_p(t, q1, q2, v1, v2, v3, fields_solver, params):
    """advection_terms.py:147-188 (calls get_fields once, which toggles)."""
    e = params.charge
    m = params.mass
    E1, E2, E3, B1, B2, B3 = fields_solver.get_fields()
    C_p1 = (e / m) * (E1 + v2 * B3 - v3 * B2)
    C_p2 = (e / m) * (E2 + v3 * B1 - v1 * B3)
    C_p3 = (e / m) * (E3 + v1 * B2 - v2 * B1)
    return C_p1, C_p2, C_p3


A_p = C_p


def integral_over_v(array, integral_measure):
    """src/utils/integral_over_v.py:6-7: ``af.sum(array, 0) * measure``."""
    return array.sum(dim=0, keepdim=True) * integral_measure


class moments:
    """src/nonrelativistic_boltzmann/moments.py:5-57."""

    @staticmethod
    def density(f, v1, v2, v3, m):
        return integral_over_v(f, m)

    @staticmethod
    def mom_v1_bulk(f, v1, v2, v3, m):
        return integral_over_v(f * v1, m)

    @staticmethod
    def mom_v2_bulk(f, v1, v2, v3, m):
        return integral_over_v(f * v2, m)

    @staticmethod
    def mom_v3_bulk(f, v1, v2, v3, m):
        return integral_over_v(f * v3, m)

    @staticmethod
    def energy(f, v1, v2, v3, m):
        return integral_over_v(0.5 * f * (v1 ** 2 + v2 ** 2 + v3 ** 2), m)

    @staticmethod
    def energy_q1(f, v1, v2, v3, m):
        return integral_over_v(0.5 * f * v1 ** 2, m)

    @staticmethod
    def energy_q2(f, v1, v2, v3, m):
        return integral_over_v(0.5 * f * v2 ** 2, m)

    @staticmethod
    def energy_q3(f, v1, v2, v3, m):
        return integral_over_v(0.5 * f * v3 ** 2, m)

    @staticmethod
    def q_q1(f, v1, v2, v3, m):
        return integral_over_v(f * v1 * (v1 ** 2 + v2 ** 2 + v3 ** 2), m)

    @staticmethod
    def q_q2(f, v1, v2, v3, m):
        return integral_over_v(f * v2 * (v1 ** 2 + v2 ** 2 + v3 ** 2), m)

    @staticmethod
    def q_q3(f, v1, v2, v3, m):
        return integral_over_v(f * v3 * (v1 ** 2 + v2 ** 2 + v3 ** 2), m)


def f0(v1, v2, v3, n, T, v1_bulk, v2_bulk, v3_bulk, params):
    """collision_operator.py:15-68: local Maxwellian for p_dim = 1, 2, 3."""
    m = params.mass
    k = params.boltzmann_constant
    if params.p_dim == 3:
        return (n * (m / (2 * math.pi * k * T)) ** (3 / 2)
                * torch.exp(-m * (v1 - v1_bulk) ** 2 / (2 * k * T))
                * torch.exp(-m * (v2 - v2_bulk) ** 2 / (2 * k * T))
                * torch.exp(-m * (v3 - v3_bulk) ** 2 / (2 * k * T)))
    if params.p_dim == 2:
        return (n * (m / (2 * math.pi * k * T))
                * torch.exp(-m * (v1 - v1_bulk) ** 2 / (2 * k * T))
                * torch.exp(-m * (v2 - v2_bulk) ** 2 / (2 * k * T)))
    return (n * torch.sqrt(m / (2 * math.pi * k * T))
            * torch.exp(-m * (v1 - v1_bulk) ** 2 / (2 * k * T)))


BGK_EPS = 1e-30


def BGK_moments(n, mom1, mom2, mom3, E, params):
    """collision_operator.py:102-116: bulk velocity and temperature."""
    m = params.mass
    eps = BGK_EPS
    v1_bulk = mom1 / (n + eps)
    v2_bulk = mom2 / (n + eps)
    v3_bulk = mom3 / (n + eps)
    T = ((1 / params.p_dim) * (2 * (E * m) - (n * m) * v1_bulk ** 2
                               - (n * m) * v2_bulk ** 2 - (n * m) * v3_bulk ** 2)
         / (n + eps) + eps)
    return v1_bulk, v2_bulk, v3_bulk, T


def BGK(f, t, q1, q2, v1, v2, v3, moments_fn, params, flag=False):
    """collision_operator.py:70-128: ``-(f - f_MB) / tau``."""
    n = moments_fn("density", f)
    v1_bulk, v2_bulk, v3_bulk, T = BGK_moments(
        n, moments_fn("mom_v1_bulk", f), moments_fn("mom_v2_bulk", f),
        moments_fn("mom_v3_bulk", f), moments_fn("energy", f), params)
    f_MB = f0(v1, v2, v3, n, T, v1_bulk, v2_bulk, v3_bulk, params)
    tau = as4(params.tau(q1, q2, v1, v2, v3))
    if not flag:
        return -(f - f_MB) / tau
    return f_MB


# ---------------------------------------------------------------------------
# Fields: nonlinear/fields/fields.py + electrodynamic/fdtd_explicit.py
# ---------------------------------------------------------------------------
def fft_poisson(rho_int, dq1, dq2, eps):
    """electrostatic/fft.py:9-78 on the interior charge density (N_q1, N_q2):
    remove the mean, phi_hat = rho_hat / (4 pi^2 |k|^2) with k from
    fftfreq(N, dq) (phi_hat[0, 0] = 0), E = -grad phi spectrally, divided by
    eps.  Returns (E1, E2) shaped (N_q1, N_q2)."""
    N1, N2 = rho_int.shape
    k1 = torch.fft.fftfreq(N1, dq1, dtype=F64)
    k2 = torch.fft.fftfreq(N2, dq2, dtype=F64)
    k1, k2 = torch.meshgrid(k1, k2, indexing="ij")
    rho = rho_int - rho_int.mean()
    rho_hat = torch.fft.fft2(rho)
    den = 4 * math.pi ** 2 * (k1 ** 2 + k2 ** 2)
    den[0, 0] = 1.0
    phi_hat = rho_hat / den
    phi_hat[0, 0] = 0
    E1 = torch.fft.ifft2(-1j * 2 * math.pi * k1 * phi_hat).real
    E2 = torch.fft.ifft2(-1j * 2 * math.pi * k2 * phi_hat).real
    return E1 / eps, E2 / eps


class OracleFields:
    """fields.py:19-690 for the electrodynamic (FDTD) path with periodic q.

    Arrays are (6, 1, Q1g, Q2g).  Reads of a component make a copy (ArrayFire
    indexing returns a new array); writes of a component mutate the shared
    object, so the aliasing of fields.py:493-494 (``at_n`` and
    ``at_n_plus_half`` bound to the same object as ``cell_centered_EM_fields``)
    is reproduced exactly."""

    def __init__(self, solver, rho_initial):
        p = solver.params
        self.params = p
        self.N_g1, self.N_g2 = solver.N_g1, solver.N_g2
        self.N_q1, self.N_q2 = solver.N_q1, solver.N_q2
        self.dq1, self.dq2 = solver.dq1, solver.dq2
        self.q1_left_bot, self.q2_left_bot = solver.q1_left_bot, solver.q2_left_bot
        self.q1_center_bot, self.q2_center_bot = solver.q1_center_bot, solver.q2_center_bot
        self.q1_left_center, self.q2_left_center = solver.q1_left_center, solver.q2_left_center
        self.q1_center, self.q2_center = solver.q1_center, solver.q2_center
        self.initialize = solver.initialize
        self.time_elapsed = 0
        self.at_n = True  # fields.py:273
        shape = (6, 1) + tuple(solver.q1_center.shape[2:])
        self.cell_centered_EM_fields = torch.zeros(shape, dtype=F64)
        self.cell_centered_EM_fields_at_n = torch.zeros(shape, dtype=F64)
        self.cell_centered_EM_fields_at_n_plus_half = torch.zeros(shape, dtype=F64)
        self.yee_grid_EM_fields = torch.zeros(shape, dtype=F64)
        if p.fields_initialize == "user-defined":  # fields.py:474-477
            self.initialize_electric_fields()
            self.initialize_magnetic_fields()
            self.yee_grid_to_cell_centered_grid()
        elif p.fields_initialize == "fft":  # fields.py:465-472
            self.compute_electrostatic_fields(rho_initial)
            self.cell_centered_grid_to_yee_grid("E")
        elif p.fields_initialize == "fft + user-defined magnetic fields":  # fields.py:479-488
            self.compute_electrostatic_fields(rho_initial)
            self.cell_centered_grid_to_yee_grid("E")
            self.initialize_magnetic_fields()
            self.yee_grid_to_cell_centered_grid("B")
        else:
            raise NotImplementedError("Method not valid/not implemented")
        # How the electrodynamic fields follow the RK stages: the reference's
        # toggle (RK2, the only stepper the reference has), or one leapfrog
        # per step driven by the stepper (SSP-RK3, timestep_fvm_ssprk3).
        rk3 = getattr(p, "time_stepper_fvm", "RK2") in ("SSP-RK3", "RK3")
        self.coupling = ("once-per-step" if rk3 and p.fields_type == "electrodynamic"
                         else "toggle")
        self.stage_fields = None
        if self.coupling == "toggle":
            # fields.py:493-494: Python aliasing, not copies.
            self.cell_centered_EM_fields_at_n = self.cell_centered_EM_fields
            self.cell_centered_EM_fields_at_n_plus_half = self.cell_centered_EM_fields
        else:
            # separate arrays; at t = 0, B^{-1/2} = B^{1/2} (fields.py:491)
            self.cell_centered_EM_fields_at_n = self.cell_centered_EM_fields.clone()
            self.cell_centered_EM_fields_at_n_plus_half = self.cell_centered_EM_fields.clone()

    def _bc(self, x):
        """communicate_fields + periodic apply_bcs_fields: refill the ghosts
        of a (6,1,Q1g,Q2g) array from the periodic interior."""
        return periodic_fill(x, self.N_g1, self.N_g2)

    def initialize_electric_fields(self):
        """fields.py:280-298 (initialize_E branch)."""
        E1 = as4(self.initialize.initialize_E(self.q1_left_center, self.q2_left_center, self.params)[0])
        E2 = as4(self.initialize.initialize_E(self.q1_center_bot, self.q2_center_bot, self.params)[1])
        E3 = as4(self.initialize.initialize_E(self.q1_center, self.q2_center, self.params)[2])
        sh = self.yee_grid_EM_fields.shape[1:]
        self.yee_grid_EM_fields[0] = torch.broadcast_to(E1, (1,) + sh)[0]
        self.yee_grid_EM_fields[1] = torch.broadcast_to(E2, (1,) + sh)[0]
        self.yee_grid_EM_fields[2] = torch.broadcast_to(E3, (1,) + sh)[0]

    def initialize_magnetic_fields(self):
        """fields.py:343-360 (initialize_B branch)."""
        B1 = as4(self.initialize.initialize_B(self.q1_center_bot, self.q2_center_bot, self.params)[0])
        B2 = as4(self.initialize.initialize_B(self.q1_left_center, self.q2_left_center, self.params)[1])
        B3 = as4(self.initialize.initialize_B(self.q1_left_bot, self.q2_left_bot, self.params)[2])
        sh = self.yee_grid_EM_fields.shape[1:]
        self.yee_grid_EM_fields[3] = torch.broadcast_to(B1, (1,) + sh)[0]
        self.yee_grid_EM_fields[4] = torch.broadcast_to(B2, (1,) + sh)[0]
        self.yee_grid_EM_fields[5] = torch.broadcast_to(B3, (1,) + sh)[0]

    def yee_grid_to_cell_centered_grid(self, which=None):
        """fields.py:527-557 (``which`` = 'E', 'B' or both)."""
        y = self.yee_grid_EM_fields.clone()
        E1, E2, E3, B1, B2, B3 = (y[i:i + 1] for i in range(6))
        cc = self.cell_centered_EM_fields
        if which in ("E", None):
            cc[0] = (0.5 * (E1 + af_shift(E1, 0, 0, -1, 0)))[0]
            cc[1] = (0.5 * (E2 + af_shift(E2, 0, 0, 0, -1)))[0]
            cc[2] = E3[0]
        if which in ("B", None):
            cc[3] = (0.5 * (B1 + af_shift(B1, 0, 0, 0, -1)))[0]
            cc[4] = (0.5 * (B2 + af_shift(B2, 0, 0, -1, 0)))[0]
            cc[5] = (0.25 * (B3 + af_shift(B3, 0, 0, 0, -1) + af_shift(B3, 0, 0, -1, 0)
                             + af_shift(B3, 0, 0, -1, -1)))[0]

    def cell_centered_grid_to_yee_grid(self, which=None):
        """fields.py:497-525 (``which`` = 'E', 'B' or both)."""
        c = self.cell_centered_EM_fields.clone()
        E1, E2, E3, B1, B2, B3 = (c[i:i + 1] for i in range(6))
        y = self.yee_grid_EM_fields
        if which in ("E", None):
            y[0] = (0.5 * (E1 + af_shift(E1, 0, 0, 1, 0)))[0]
            y[1] = (0.5 * (E2 + af_shift(E2, 0, 0, 0, 1)))[0]
            y[2] = E3[0]
        if which in ("B", None):
            y[3] = (0.5 * (B1 + af_shift(B1, 0, 0, 0, 1)))[0]
            y[4] = (0.5 * (B2 + af_shift(B2, 0, 0, 1, 0)))[0]
            y[5] = (0.25 * (B3 + af_shift(B3, 0, 0, 1, 0) + af_shift(B3, 0, 0, 0, 1)
                            + af_shift(B3, 0, 0, 1, 1)))[0]

    def compute_electrostatic_fields(self, rho):
        """fields.py:559-566 + fft.py:9-78: sum rho over species, solve on the
        interior, write E1, E2 of the cell-centred fields and refill ghosts
        (communicate_fields + periodic apply_bcs_fields)."""
        rho = rho.sum(dim=1, keepdim=True)
        g1, g2 = self.N_g1, self.N_g2
        rho_int = rho[0, 0, g1:rho.shape[2] - g1, g2:rho.shape[3] - g2]
        E1, E2 = fft_poisson(rho_int, self.dq1, self.dq2, self.params.eps)
        cc = self.cell_centered_EM_fields
        cc[0, 0, g1:cc.shape[2] - g1, g2:cc.shape[3] - g2] = E1
        cc[1, 0, g1:cc.shape[2] - g1, g2:cc.shape[3] - g2] = E2
        cc.copy_(self._bc(cc))

    def fdtd_evolve_E(self, dt):
        """fdtd_explicit.py:9-82."""
        eps, mu = self.params.eps, self.params.mu
        y = self.yee_grid_EM_fields
        B1, B2, B3 = y[3:4].clone(), y[4:5].clone(), y[5:6].clone()
        B1_plus_q2 = af_shift(B1, 0, 0, 0, -1)
        B2_plus_q1 = af_shift(B2, 0, 0, -1, 0)
        B3_plus_q1 = af_shift(B3, 0, 0, -1, 0)
        B3_plus_q2 = af_shift(B3, 0, 0, 0, -1)
        curlB_1 = (B3_plus_q2 - B3) / self.dq2
        curlB_2 = -(B3_plus_q1 - B3) / self.dq1
        curlB_3 = (B2_plus_q1 - B2) / self.dq1 - (B1_plus_q2 - B1) / self.dq2
        y[0] = (y[0:1] + ((dt / (mu * eps)) * curlB_1 - self.J1 * dt / eps))[0]
        y[1] = (y[1:2] + ((dt / (mu * eps)) * curlB_2 - self.J2 * dt / eps))[0]
        y[2] = (y[2:3] + ((dt / (mu * eps)) * curlB_3 - self.J3 * dt / eps))[0]

    def fdtd_evolve_B(self, dt):
        """fdtd_explicit.py:84-138."""
        y = self.yee_grid_EM_fields
        E1, E2, E3 = y[0:1].clone(), y[1:2].clone(), y[2:3].clone()
        E1_minus_q2 = af_shift(E1, 0, 0, 0, 1)
        E2_minus_q1 = af_shift(E2, 0, 0, 1, 0)
        E3_minus_q1 = af_shift(E3, 0, 0, 1, 0)
        E3_minus_q2 = af_shift(E3, 0, 0, 0, 1)
        curlE_1 = (E3 - E3_minus_q2) / self.dq2
        curlE_2 = -(E3 - E3_minus_q1) / self.dq1
        curlE_3 = (E2 - E2_minus_q1) / self.dq1 - (E1 - E1_minus_q2) / self.dq2
        y[3] = (y[3:4] + -dt * curlE_1)[0]
        y[4] = (y[4:5] + -dt * curlE_2)[0]
        y[5] = (y[5:6] + -dt * curlE_3)[0]

    def fdtd(self, dt):
        """fdtd_explicit.py:141-177 (ghost refresh before each half)."""
        self.yee_grid_EM_fields = self._bc(self.yee_grid_EM_fields)
        self.fdtd_evolve_E(dt)
        self.yee_grid_EM_fields = self._bc(self.yee_grid_EM_fields)
        self.fdtd_evolve_B(dt)

    def evolve_electrodynamic_fields(self, J1, J2, J3, dt):
        """fields.py:568-647 (component writes mutate the shared object)."""
        self.J1 = J1.sum(dim=1, keepdim=True)
        self.J2 = J2.sum(dim=1, keepdim=True)
        self.J3 = J3.sum(dim=1, keepdim=True)
        E_at_n = self.cell_centered_EM_fields[:3].clone()
        self.cell_centered_EM_fields_at_n[:3] = E_at_n
        B_at_n_minus_half = self.cell_centered_EM_fields_at_n_plus_half[3:].clone()
        B_at_n_plus_half = self.cell_centered_EM_fields[3:].clone()
        B_at_n = 0.5 * (B_at_n_minus_half + B_at_n_plus_half)
        self.cell_centered_EM_fields_at_n[3:] = B_at_n
        self.cell_centered_EM_fields_at_n_plus_half[3:] = B_at_n_plus_half
        self.fdtd(dt)
        self.yee_grid_to_cell_centered_grid()
        E_at_n_plus_one = self.cell_centered_EM_fields[:3].clone()
        E_at_n_plus_half = 0.5 * (E_at_n + E_at_n_plus_one)
        self.cell_centered_EM_fields_at_n_plus_half[:3] = E_at_n_plus_half
        self.time_elapsed += dt

    def evolve_once_per_step(self, J1, J2, J3, dt):
        """The once-per-step leapfrog of SSP-RK3 (timestep_fvm_ssprk3) with
        the time levels of fields.py:568-647, kept in separate arrays: before
        it the cell-centred fields hold (E^n, B^{n+1/2}); after it they hold
        (E^{n+1}, B^{n+3/2}), ``..._at_n`` holds (E^{n+1}, B^{n+1}) with
        B^{n+1} = (B^{n+1/2} + B^{n+3/2}) / 2, and ``..._at_n_plus_half``
        holds (E^{n+1/2}, B^{n+1/2}) with E^{n+1/2} = (E^n + E^{n+1}) / 2."""
        self.J1 = J1.sum(dim=1, keepdim=True)
        self.J2 = J2.sum(dim=1, keepdim=True)
        self.J3 = J3.sum(dim=1, keepdim=True)
        E_n = self.cell_centered_EM_fields[:3].clone()
        B_n_plus_half = self.cell_centered_EM_fields[3:].clone()
        self.fdtd(dt)
        self.yee_grid_to_cell_centered_grid()
        E_n_plus_one = self.cell_centered_EM_fields[:3].clone()
        B_n_plus_3half = self.cell_centered_EM_fields[3:].clone()
        self.cell_centered_EM_fields_at_n = torch.cat(
            [E_n_plus_one, 0.5 * (B_n_plus_half + B_n_plus_3half)])
        self.cell_centered_EM_fields_at_n_plus_half = torch.cat(
            [0.5 * (E_n + E_n_plus_one), B_n_plus_half])
        self.time_elapsed += dt

    def get_fields(self):
        """fields.py:649-690 (FDTD branch; toggles ``at_n`` on every call when
        the p-solver is FVM).  With the once-per-step coupling the stepper
        sets ``stage_fields`` for each stage (outside a step: the fields at
        n) and nothing toggles."""
        if self.coupling == "once-per-step":
            src = self.stage_fields
            if src is None:
                src = self.cell_centered_EM_fields_at_n
            return tuple(src[i:i + 1].clone() for i in range(6))
        if self.params.fields_solver != "fdtd":
            src = self.cell_centered_EM_fields
        else:
            src = (self.cell_centered_EM_fields_at_n if self.at_n
                   else self.cell_centered_EM_fields_at_n_plus_half)
        out = tuple(src[i:i + 1].clone() for i in range(6))
        if self.params.solver_method_in_p == "FVM":
            self.at_n = not self.at_n
        return out


# ---------------------------------------------------------------------------
# Ghost zones: communicate.py:8-38 with periodic PETSc DMDA globalToLocal
# ---------------------------------------------------------------------------
def periodic_fill(x, N_g1, N_g2):
    """Refill ghost zones along axes 2 (q1) and 3 (q2) from the periodic
    interior, as DMDA.globalToLocal does with periodic boundaries."""
    x = x.clone()
    for axis, g in ((2, N_g1), (3, N_g2)):
        if g == 0:
            continue
        n = x.shape[axis] - 2 * g
        interior = x.narrow(axis, g, n)
        idx = (torch.arange(x.shape[axis]) - g) % n
        x = interior.index_select(axis, idx)
    return x


def ghosted_fill(x, N_g1, N_g2, periodic1, periodic2):
    """communicate.py:8-38 when an axis is not periodic: the DMDA is created
    with a 'ghosted' boundary on that axis (nonlinear_solver.py:177-201,
    265-278), globalToLocal fills only the periodic axes, and the local
    vector's entries in a ghosted boundary zone are never written, so they
    stay at the zero they were created with; petsc_local_array_to_af copies
    them into f (af_petsc_conversion.py:23-41)."""
    x = periodic_fill(x, N_g1 if periodic1 else 0, N_g2 if periodic2 else 0)
    x = x.clone()
    if not periodic1 and N_g1:
        x[:, :, :N_g1] = 0
        x[:, :, x.shape[2] - N_g1:] = 0
    if not periodic2 and N_g2:
        x[:, :, :, :N_g2] = 0
        x[:, :, :, x.shape[3] - N_g2:] = 0
    return x


# ---------------------------------------------------------------------------
# Solver: nonlinear_solver.py + finite_volume/{df_dt_fvm,fvm_operator}.py
# ---------------------------------------------------------------------------
class OracleSolver:
    """nonlinear_solver.py:55-521 for the FVM path (periodic, none, mirror,
    dirichlet and mirror+dirichlet q-BCs on a cartesian p-grid, one rank).  ``system`` is any object with the attributes of
    physical_system (domain sizes, params, initial_conditions, C_q, C_p,
    source, moments, boundary_conditions)."""

    def __init__(self, system):
        self.physical_system = system
        p = self.params = system.params
        self.N_q1, self.N_q2 = system.N_q1, system.N_q2
        self.N_p1, self.N_p2, self.N_p3 = system.N_p1, system.N_p2, system.N_p3
        self.dq1, self.dq2 = system.dq1, system.dq2
        self.N_species = len(p.charge)
        bc = system.boundary_conditions
        self.N_ghost = system.N_ghost
        self.N_g1 = 0 if bc.in_q1_left == "none" else self.N_ghost  # :197-201
        self.N_g2 = 0 if bc.in_q2_bottom == "none" else self.N_ghost
        for side in ("in_q1_left", "in_q1_right", "in_q2_bottom", "in_q2_top"):
            if getattr(bc, side) not in ("periodic", "none", "mirror", "dirichlet",
                                         "mirror+dirichlet"):
                raise NotImplementedError("oracle restates periodic, none, mirror and "
                                          "dirichlet q-BCs")
        # nonlinear_solver.py:187-195: periodic DMDA axes, else 'ghosted'
        self.periodic1 = bc.in_q1_left == "periodic"
        self.periodic2 = bc.in_q2_bottom == "periodic"
        self.boundary_conditions = bc
        # nonlinear_solver.py:133-145: species lists -> (1, S) arrays
        S = self.N_species
        self.mass = torch.tensor(list(p.mass), dtype=F64).reshape(1, S, 1, 1)
        self.charge = torch.tensor(list(p.charge), dtype=F64).reshape(1, S, 1, 1)
        self.params = types.SimpleNamespace(**{k: getattr(p, k) for k in dir(p)
                                               if not k.startswith("__")})
        self.params.mass = self.mass
        self.params.charge = self.charge
        (qlb, qcb, qlc, qc) = calculate_q(system.q1_start, system.q2_start,
                                          self.N_q1, self.N_q2, self.N_g1, self.N_g2,
                                          self.dq1, self.dq2)
        self.q1_left_bot, self.q2_left_bot = qlb
        self.q1_center_bot, self.q2_center_bot = qcb
        self.q1_left_center, self.q2_left_center = qlc
        self.q1_center, self.q2_center = qc
        dp1, dp2, dp3 = system.dp1, system.dp2, system.dp3
        self.p1_center, self.p2_center, self.p3_center = calculate_p_center(
            system.p1_start, system.p2_start, system.p3_start,
            self.N_p1, self.N_p2, self.N_p3, dp1, dp2, dp3)
        self.p1_left, self.p2_bottom, self.p3_back = calculate_p_corner(
            system.p1_start, system.p2_start, system.p3_start,
            self.N_p1, self.N_p2, self.N_p3, dp1, dp2, dp3)
        self.dp1 = torch.tensor(dp1, dtype=F64).reshape(1, S, 1, 1)  # :378-380
        self.dp2 = torch.tensor(dp2, dtype=F64).reshape(1, S, 1, 1)
        self.dp3 = torch.tensor(dp3, dtype=F64).reshape(1, S, 1, 1)
        self.p2_left, self.p3_left = self.p2_center, self.p3_center  # :391-398
        self.p1_bottom, self.p3_bottom = self.p1_center, self.p3_center
        self.p1_back, self.p2_back = self.p1_center, self.p2_center
        self.initialize = system.initial_conditions
        self._C_q = system.C_q
        self._C_p = system.C_p
        self._source = system.source
        self.moments_module = system.moments
        self.time_elapsed = 0
        self.dt = None
        # nonlinear_solver.py:464-500
        self.f = as4(self.initialize.initialize_f(
            self.q1_center, self.q2_center, self.p1_center, self.p2_center,
            self.p3_center, self.params))
        self.f = torch.broadcast_to(
            self.f, (self.N_p1 * self.N_p2 * self.N_p3, S) + tuple(self.q1_center.shape[2:])).clone()
        self.fields_solver = None
        if p.fields_enabled:
            rho_initial = self.charge * self.compute_moments("density")
            self.fields_solver = OracleFields(self, rho_initial)

    # -- layout helpers: nonlinear_solver.py:416-462
    def _convert_to_q_expanded(self, x):
        S = self.N_species
        return moddims(x, self.N_p1 * self.N_p2 * self.N_p3, S,
                       self.f.shape[2], self.f.shape[3])

    def _convert_to_p_expanded(self, x):
        S = self.N_species
        return moddims(x, self.N_p1, self.N_p2, self.N_p3,
                       S * self.f.shape[2] * self.f.shape[3])

    def _communicate_f(self):
        """communicate.py:8-38 then apply_bcs_f (boundaries.py:415-577), the
        ghost refresh before every RHS evaluation (fvm_operator.py:22-33)."""
        self.f = ghosted_fill(self.f, self.N_g1, self.N_g2, self.periodic1, self.periodic2)
        self.apply_bcs_f()

    # -- boundaries.py ------------------------------------------------------
    def _mirror(self, boundary):
        """apply_mirror_bcs_f_cartesian (boundaries.py:227-311): the ghost
        zone is the mirror image of the first N_g interior cells, with the
        velocity component normal to the wall reversed (p1 for left/right,
        p2 for bottom/top)."""
        g = self.N_ghost
        f = self.f.clone()
        if boundary == "left":
            f[:, :, :g] = torch.flip(f[:, :, g:2 * g], (2,))
            f[:, :, :g] = self._convert_to_q_expanded(
                torch.flip(self._convert_to_p_expanded(f), (0,)))[:, :, :g]
        elif boundary == "right":
            n = f.shape[2]
            f[:, :, n - g:] = torch.flip(f[:, :, n - 2 * g:n - g], (2,))
            f[:, :, n - g:] = self._convert_to_q_expanded(
                torch.flip(self._convert_to_p_expanded(f), (0,)))[:, :, n - g:]
        elif boundary == "bottom":
            f[:, :, :, :g] = torch.flip(f[:, :, :, g:2 * g], (3,))
            f[:, :, :, :g] = self._convert_to_q_expanded(
                torch.flip(self._convert_to_p_expanded(f), (1,)))[:, :, :, :g]
        else:
            n = f.shape[3]
            f[:, :, :, n - g:] = torch.flip(f[:, :, :, n - 2 * g:n - g], (3,))
            f[:, :, :, n - g:] = self._convert_to_q_expanded(
                torch.flip(self._convert_to_p_expanded(f), (1,)))[:, :, :, n - g:]
        self.f = f

    def _dirichlet(self, boundary):
        """apply_dirichlet_bcs_f (boundaries.py:147-225): the user's
        f_<boundary>(f, t, q1, q2, p1, p2, p3, params) replaces the ghost
        zone where the characteristic enters the domain (C_q1 > 0 on the left,
        < 0 on the right; C_q2 likewise on bottom / top); elsewhere the ghost
        zone keeps what it holds."""
        g = self.N_ghost
        p = self.params
        C_q1, C_q2 = self._C_q(self.time_elapsed, self.q1_center, self.q2_center,
                               self.p1_center, self.p2_center, self.p3_center, p)
        full = self.f.shape
        C_q1 = torch.broadcast_to(as4(C_q1), full)
        C_q2 = torch.broadcast_to(as4(C_q2), full)
        bcm = self.boundary_conditions
        args = (self.f, self.time_elapsed, self.q1_center, self.q2_center,
                self.p1_center, self.p2_center, self.p3_center, p)
        f = self.f.clone()
        if boundary == "left":
            val = torch.where(C_q1 > 0, torch.broadcast_to(as4(bcm.f_left(*args)), full), f)
            f[:, :, :g] = val[:, :, :g]
        elif boundary == "right":
            n = full[2]
            val = torch.where(C_q1 < 0, torch.broadcast_to(as4(bcm.f_right(*args)), full), f)
            f[:, :, n - g:] = val[:, :, n - g:]
        elif boundary == "bottom":
            val = torch.where(C_q2 > 0, torch.broadcast_to(as4(bcm.f_bottom(*args)), full), f)
            f[:, :, :, :g] = val[:, :, :, :g]
        else:
            n = full[3]
            val = torch.where(C_q2 < 0, torch.broadcast_to(as4(bcm.f_top(*args)), full), f)
            f[:, :, :, n - g:] = val[:, :, :, n - g:]
        self.f = f

    def apply_bcs_f(self):
        """boundaries.py:415-577 on one rank: left, right, bottom, top in
        that order; periodic and none need nothing; 'mirror+dirichlet'
        mirrors first.  Cartesian p-grid only (p_space_grid, :437-450)."""
        bc = self.boundary_conditions
        for boundary, name in (("left", "in_q1_left"), ("right", "in_q1_right"),
                               ("bottom", "in_q2_bottom"), ("top", "in_q2_top")):
            kind = getattr(bc, name)
            if kind in ("periodic", "none"):
                continue
            if kind in ("mirror", "mirror+dirichlet"):
                if getattr(self.params, "p_space_grid", "cartesian") != "cartesian":
                    raise NotImplementedError("Unsupported coordinate system in p_space")
                self._mirror(boundary)
            if kind in ("dirichlet", "mirror+dirichlet"):
                self._dirichlet(boundary)

    def compute_moments(self, moment_name, f=None):
        """compute_moments.py:7-59 (cartesian p-grid)."""
        if f is None:
            f = self.f
        fn = getattr(self.moments_module, moment_name)
        return fn(f, self.p1_center, self.p2_center, self.p3_center,
                  self.dp3 * self.dp2 * self.dp1)

    def _q_faces(self, f):
        """df_dt_fvm.py:54-80: reconstruction along q1 / q2, the face states
        either side of every left / bottom face, the face velocities C_q and
        the upwind states f_left / f_bot (riemann.py:8-58)."""
        p = self.params
        rq = p.reconstruction_method_in_q
        f_left_plus_eps, f_right_minus_eps = reconstruct(f, 2, rq)
        f_bot_plus_eps, f_top_minus_eps = reconstruct(f, 3, rq)
        f_left_minus_eps = af_shift(f_right_minus_eps, 0, 0, 1)
        f_bot_minus_eps = af_shift(f_top_minus_eps, 0, 0, 0, 1)
        C_q1 = as4(self._C_q(self.time_elapsed, self.q1_left_center, self.q2_left_center,
                             self.p1_center, self.p2_center, self.p3_center, p)[0])
        C_q2 = as4(self._C_q(self.time_elapsed, self.q1_center_bot, self.q2_center_bot,
                             self.p1_center, self.p2_center, self.p3_center, p)[1])
        f_left = riemann_upwind_state(f_left_minus_eps, f_left_plus_eps, C_q1)
        f_bot = riemann_upwind_state(f_bot_minus_eps, f_bot_plus_eps, C_q2)
        return (f_left_minus_eps, f_left_plus_eps, f_bot_minus_eps, f_bot_plus_eps,
                C_q1, C_q2, f_left, f_bot)

    def currents(self, f):
        """J1, J2, J3 of df_dt_fvm.py:214-224: the charge times the bulk
        momenta of the upwind face states f_left (J1, at (i, j+1/2)) and f_bot
        (J2, at (i+1/2, j)) and of f itself (J3)."""
        f_left, f_bot = self._q_faces(f)[6:]
        J1 = self.charge * self.compute_moments("mom_v1_bulk", f=f_left)
        J2 = self.charge * self.compute_moments("mom_v2_bulk", f=f_bot)
        J3 = self.charge * self.compute_moments("mom_v3_bulk", f=f)
        return J1, J2, J3

    def df_dt_fvm(self, f):
        """df_dt_fvm.py:33-339 (the hybrid-model and electrostatic branches are
        not restated)."""
        p = self.params
        df_dt = 0
        f_left = f_bot = None
        if p.solver_method_in_q == "FVM":
            (f_left_minus_eps, f_left_plus_eps, f_bot_minus_eps, f_bot_plus_eps,
             C_q1, C_q2, f_left, f_bot) = self._q_faces(f)
            rs = getattr(p, "riemann_solver_in_q", "upwind-flux")
            left_flux = face_flux(rs, f_left_minus_eps, f_left_plus_eps, C_q1)
            bot_flux = face_flux(rs, f_bot_minus_eps, f_bot_plus_eps, C_q2)
            right_flux = af_shift(left_flux, 0, 0, -1)
            top_flux = af_shift(bot_flux, 0, 0, 0, -1)
            df_dt = df_dt + (-(right_flux - left_flux) / self.dq1
                             - (top_flux - bot_flux) / self.dq2)
            if p.source_enabled:
                df_dt = df_dt + self._source(f, self.time_elapsed, self.q1_center,
                                             self.q2_center, self.p1_center,
                                             self.p2_center, self.p3_center,
                                             self.compute_moments, p, False)
        if p.solver_method_in_p == "FVM" and p.fields_enabled:
            fs = self.fields_solver
            # reference coupling (RK2): FDTD inside the RHS whenever at_n is
            # False; the once-per-step coupling of SSP-RK3 advances the fields
            # from the stepper instead
            if p.fields_type == "electrodynamic" and fs.coupling == "toggle" \
                    and fs.at_n is False:
                J1 = self.charge * self.compute_moments("mom_v1_bulk", f=f_left)
                J2 = self.charge * self.compute_moments("mom_v2_bulk", f=f_bot)
                J3 = self.charge * self.compute_moments("mom_v3_bulk", f=f)
                fs.evolve_electrodynamic_fields(J1, J2, J3, self.dt)
            if p.fields_type == "electrostatic":  # df_dt_fvm.py:235-242
                if p.fields_solver != "fft":
                    raise NotImplementedError("only the FFT electrostatic solver is restated")
                rho = self.charge * self.compute_moments("density", f=f)
                fs.compute_electrostatic_fields(rho)
            C_p1 = as4(self._C_p(self.time_elapsed, self.q1_center, self.q2_center,
                                 self.p1_left, self.p2_left, self.p3_left, fs, p)[0])
            C_p2 = as4(self._C_p(self.time_elapsed, self.q1_center, self.q2_center,
                                 self.p1_bottom, self.p2_bottom, self.p3_bottom, fs, p)[1])
            C_p3 = as4(self._C_p(self.time_elapsed, self.q1_center, self.q2_center,
                                 self.p1_back, self.p2_back, self.p3_back, fs, p)[2])
            full = f.shape
            C_p1 = self._convert_to_p_expanded(torch.broadcast_to(C_p1, full))
            C_p2 = self._convert_to_p_expanded(torch.broadcast_to(C_p2, full))
            C_p3 = self._convert_to_p_expanded(torch.broadcast_to(C_p3, full))
            fp = self._convert_to_p_expanded(f)
            rp = p.reconstruction_method_in_p
            f_left_plus_eps, f_right_minus_eps = reconstruct(fp, 0, rp)
            f_bot_plus_eps, f_top_minus_eps = reconstruct(fp, 1, rp)
            f_back_plus_eps, f_front_minus_eps = reconstruct(fp, 2, rp)
            f_left_minus_eps = af_shift(f_right_minus_eps, 1)
            f_bot_minus_eps = af_shift(f_top_minus_eps, 0, 1)
            f_back_minus_eps = af_shift(f_front_minus_eps, 0, 0, 1)
            rs = getattr(p, "riemann_solver_in_p", "upwind-flux")
            left_flux_p1 = face_flux(rs, f_left_minus_eps, f_left_plus_eps, C_p1)
            bot_flux_p2 = face_flux(rs, f_bot_minus_eps, f_bot_plus_eps, C_p2)
            back_flux_p3 = face_flux(rs, f_back_minus_eps, f_back_plus_eps, C_p3)
            right_flux_p1 = af_shift(left_flux_p1, -1)
            top_flux_p2 = af_shift(bot_flux_p2, 0, -1)
            front_flux_p3 = af_shift(back_flux_p3, 0, 0, -1)
            q = self._convert_to_q_expanded
            d1 = (q(right_flux_p1) - q(left_flux_p1)) * (1 / self.dp1)
            d2 = (q(top_flux_p2) - q(bot_flux_p2)) * (1 / self.dp2)
            d3 = (q(front_flux_p3) - q(back_flux_p3)) * (1 / self.dp3)
            df_dt = df_dt - (d1 + d2 + d3)
        return df_dt

    # -- time steppers ------------------------------------------------------
    def timestep_fvm_rk2(self, dt):
        """fvm_operator.py:6-35 (RK2 midpoint, ghost refresh before each RHS)."""
        self._communicate_f()
        f_initial = self.f
        self.f = self.f + self.df_dt_fvm(self.f) * (dt / 2)
        self._communicate_f()
        self.f = f_initial + self.df_dt_fvm(self.f) * dt

    def timestep_fvm_ssprk3(self, dt):
        """SSP-RK3 (Shu-Osher); added, the reference only has RK2 midpoint.
        u1 = u + dt L(u); u2 = 3/4 u + 1/4 (u1 + dt L(u1));
        u^{n+1} = 1/3 u + 2/3 (u2 + dt L(u2)).

        Electrodynamic fields (the reference defines no RK3 coupling; its
        per-C_p toggle, fields.py:685-688, would run FDTD on 1, 2, 1, 2, ...
        of the stages and advance the fields 1.5 dt per step): the stages sit
        at t^n, t^{n+1} and t^{n+1/2}, so stage 1 sees the fields at n, then
        one FDTD leapfrog with J^{n+1/2} = (J(u^n) + J(u1)) / 2 (u1 is the
        Euler estimate of u^{n+1}) takes E^n -> E^{n+1} and B^{n+1/2} ->
        B^{n+3/2}, stage 2 sees the fields at n+1 and stage 3 those at
        n+1/2."""
        fs = self.fields_solver
        once = fs is not None and fs.coupling == "once-per-step"
        self._communicate_f()
        un = self.f
        if once:
            J_n = self.currents(un)
            fs.stage_fields = fs.cell_centered_EM_fields_at_n
        u1 = un + self.df_dt_fvm(un) * dt
        self.f = u1
        self._communicate_f()
        u1 = self.f
        if once:
            J_1 = self.currents(u1)
            fs.evolve_once_per_step(*(0.5 * (a + b) for a, b in zip(J_n, J_1)), dt)
            fs.stage_fields = fs.cell_centered_EM_fields_at_n
        u2 = 0.75 * un + 0.25 * (u1 + self.df_dt_fvm(u1) * dt)
        self.f = u2
        self._communicate_f()
        u2 = self.f
        if once:
            fs.stage_fields = fs.cell_centered_EM_fields_at_n_plus_half
        self.f = (1 / 3) * un + (2 / 3) * (u2 + self.df_dt_fvm(u2) * dt)
        if once:
            fs.stage_fields = None

    def check_divergence(self):
        """timestep.py:28-38."""
        if bool(torch.isinf(self.f).any()) or bool(torch.isnan(self.f).any()):
            raise SystemExit("Solver Diverging!")

    def strang_timestep(self, dt):
        """timestep.py:101-160 (FVM in q and p: no splitting, :117-125)."""
        self.dt = dt
        stepper = getattr(self.params, "time_stepper_fvm", "RK2")
        if stepper == "RK2":
            self.timestep_fvm_rk2(dt)
        elif stepper in ("SSP-RK3", "RK3"):
            self.timestep_fvm_ssprk3(dt)
        else:
            raise NotImplementedError(stepper)
        self.check_divergence()
        self.time_elapsed += dt

    lie_timestep = swss_timestep = jia_timestep = strang_timestep

    # -- convenience for tests ---------------------------------------------
    def interior(self, x=None):
        x = self.f if x is None else x
        g1, g2 = self.N_g1, self.N_g2
        return x[:, :, g1:x.shape[2] - g1, g2:x.shape[3] - g2]


def make_system(domain, boundary_conditions, params, initial_conditions,
                advection_term, source, moments_module):
    """Attribute container equivalent to physical_system (physical_system.py:
    147-213): grid extents, dq = L/N, per-species dp lists and callbacks."""
    s = types.SimpleNamespace()
    for k in ("N_q1", "q1_start", "q1_end", "N_q2", "q2_start", "q2_end",
              "N_p1", "p1_start", "p1_end", "N_p2", "p2_start", "p2_end",
              "N_p3", "p3_start", "p3_end", "N_ghost"):
        setattr(s, k, getattr(domain, k))
    s.dq1 = (s.q1_end - s.q1_start) / s.N_q1
    s.dq2 = (s.q2_end - s.q2_start) / s.N_q2
    n = len(params.charge)
    s.dp1 = [(s.p1_end[i] - s.p1_start[i]) / s.N_p1 for i in range(n)]
    s.dp2 = [(s.p2_end[i] - s.p2_start[i]) / s.N_p2 for i in range(n)]
    s.dp3 = [(s.p3_end[i] - s.p3_start[i]) / s.N_p3 for i in range(n)]
    s.boundary_conditions = boundary_conditions
    s.params = params
    s.initial_conditions = initial_conditions
    s.C_q = advection_term.C_q
    s.C_p = advection_term.C_p
    s.source = source
    s.moments = moments_module
    return s


class _OraclePlugins:
    """The oracle's restatement of the built-in plugins, shaped like the
    reference's advection_terms module."""
    C_q = staticmethod(C_q)
    A_q = staticmethod(A_q)
    C_p = staticmethod(C_p)
    A_p = staticmethod(A_p)


advection_terms = _OraclePlugins
