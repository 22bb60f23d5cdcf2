"""Electrodynamic fields for the FVM step: mirror of
bolt/lib/nonlinear/fields/fields.py:19-690 (FDTD branch) on the GPU.

Field arrays are FP64 [6][Q] device tensors (E1, E2, E3, B1, B2, B3; Q interior
q-cells, periodic, no ghost zones).  The FDTD update and the Yee-to-centre
averaging run in the HIP library.  Two couplings to the RK stages:

- 'toggle' (RK2, the reference's stepper): the staggered bookkeeping of
  evolve_electrodynamic_fields (fields.py:568-647) as in the reference, run
  from the RHS whenever at_n is False, with at_n flipped by every C_p
  evaluation (fields.py:685-688), and fields.py:493-494, which binds
  ``cell_centered_EM_fields_at_n`` and ``..._at_n_plus_half`` to the *same*
  array as ``cell_centered_EM_fields`` (component writes through any of the
  three names change all of them);
- 'once-per-step' (SSP-RK3, which the reference does not have: its toggle
  would run FDTD on 1, 2, 1, 2, ... of the stages and advance the fields
  1.5 dt per step): the stages sit at t^n, t^{n+1}, t^{n+1/2} and see the
  fields at n, n+1 and n+1/2; one leapfrog per step, between stages 1 and 2,
  with J^{n+1/2} = (J(u^n) + J(u1)) / 2, and separate arrays per time level.
"""
import ctypes
import math

import torch

from . import _lib

F64 = torch.float64


class fields_solver(object):
    """fields.py:19-690 for fields_type 'electrodynamic' with fields_solver
    'fdtd', and fields_type 'electrostatic' with fields_solver 'fft'."""

    def __init__(self, solver):
        self._solver = solver
        p = self.params = solver.params
        self.N_q1, self.N_q2 = solver.N_q1, solver.N_q2
        self.dq1, self.dq2 = solver.dq1, solver.dq2
        self.time_elapsed = 0
        self.at_n = True  # fields.py:273
        Q = solver.N_q1 * solver.N_q2
        dev = solver.device
        self._yee = torch.zeros(6, Q, dtype=F64, device=dev)
        cc = torch.zeros(6, Q, dtype=F64, device=dev)
        self._fdtd = getattr(p, 'fields_solver', 'fdtd').lower() == 'fdtd'
        init = p.fields_initialize
        if init == 'user-defined':  # fields.py:474-477
            self._initialize_electric_fields()
            self._initialize_magnetic_fields()
            self._yee_to_cc(cc)
        elif init in ('fft', 'fft + user-defined magnetic fields'):  # fields.py:465-488
            rho = solver._rho()
            self._poisson(rho, cc)
            self._cc_to_yee_E(cc)
            if init != 'fft':
                self._initialize_magnetic_fields()
                tmp = torch.empty_like(cc)
                self._yee_to_cc(tmp)
                cc[3:] = tmp[3:]
        else:
            raise NotImplementedError('Method not valid/not implemented')
        self._cc = cc
        rk3 = getattr(p, 'time_stepper_fvm', 'RK2') == 'SSP-RK3'
        self.coupling = 'once-per-step' if rk3 and self._fdtd else 'toggle'
        self.stage_fields = None
        if self.coupling == 'toggle':
            # fields.py:493-494: all three names refer to one object.
            self._cc_at_n = cc
            self._cc_at_n_plus_half = cc
        else:
            # separate time levels; at t = 0, B^{-1/2} = B^{1/2} (fields.py:491)
            self._cc_at_n = cc.clone()
            self._cc_at_n_plus_half = cc.clone()

    # -- initialisation (fields.py:280-419) ----------------------------------
    def _eval_interior(self, fn, q1, q2, k):
        s = self._solver
        val = fn(q1, q2, s.params)[k]
        val = torch.as_tensor(val, dtype=F64, device=s.device)
        val = torch.broadcast_to(val, q1.shape)
        return s._q_interior(val).reshape(-1)

    def _initialize_electric_fields(self):
        s = self._solver
        ini = s.physical_system.initial_conditions
        if not hasattr(ini, 'initialize_E'):
            raise NotImplementedError('Initialization method for electric fields not valid/found')
        self._yee[0] = self._eval_interior(ini.initialize_E, s.q1_left_center, s.q2_left_center, 0)
        self._yee[1] = self._eval_interior(ini.initialize_E, s.q1_center_bot, s.q2_center_bot, 1)
        self._yee[2] = self._eval_interior(ini.initialize_E, s.q1_center, s.q2_center, 2)

    def _initialize_magnetic_fields(self):
        s = self._solver
        ini = s.physical_system.initial_conditions
        if not hasattr(ini, 'initialize_B'):
            raise NotImplementedError('Initialization method for magnetic fields not valid/found')
        self._yee[3] = self._eval_interior(ini.initialize_B, s.q1_center_bot, s.q2_center_bot, 0)
        self._yee[4] = self._eval_interior(ini.initialize_B, s.q1_left_center, s.q2_left_center, 1)
        self._yee[5] = self._eval_interior(ini.initialize_B, s.q1_left_bot, s.q2_left_bot, 2)

    def _poisson(self, rho, cc):
        """electrostatic/fft.py:9-78 on the device (rocFFT through torch.fft):
        rho ([Q], index i1 + N_q1 i2) minus its mean, phi_hat = rho_hat /
        (4 pi^2 |k|^2) with phi_hat[0, 0] = 0, E = -grad phi / eps into rows 0
        and 1 of cc."""
        s = self._solver
        N1, N2 = s.N_q1, s.N_q2
        r = rho.reshape(N2, N1)
        r = r - r.mean()
        k1 = torch.fft.fftfreq(N1, float(self.dq1), dtype=F64, device=r.device).reshape(1, N1)
        k2 = torch.fft.fftfreq(N2, float(self.dq2), dtype=F64, device=r.device).reshape(N2, 1)
        den = 4 * math.pi ** 2 * (k1 ** 2 + k2 ** 2)
        den[0, 0] = 1.0
        phi = torch.fft.fft2(r) / den
        phi[0, 0] = 0
        eps = float(self.params.eps)
        cc[0] = (torch.fft.ifft2(-2j * math.pi * k1 * phi).real / eps).reshape(-1)
        cc[1] = (torch.fft.ifft2(-2j * math.pi * k2 * phi).real / eps).reshape(-1)

    def compute_electrostatic_fields(self, rho):
        """fields.py:559-566; rho is [Q], already summed over species."""
        self._poisson(rho, self._cc)

    def _cc_to_yee_E(self, cc):
        """fields.py:497-525 for E: E1 at (i, j+1/2) = (E1[i] + E1[i-1]) / 2,
        E2 at (i+1/2, j) = (E2[j] + E2[j-1]) / 2, E3 unchanged."""
        s = self._solver
        N1, N2 = s.N_q1, s.N_q2
        e1 = cc[0].reshape(N2, N1)
        e2 = cc[1].reshape(N2, N1)
        self._yee[0] = (0.5 * (e1 + torch.roll(e1, 1, 1))).reshape(-1)
        self._yee[1] = (0.5 * (e2 + torch.roll(e2, 1, 0))).reshape(-1)
        self._yee[2] = cc[2]

    def _yee_to_cc(self, cc):
        s = self._solver
        _lib.check(s._L.boltx_yee_to_cell_centered(ctypes.byref(s._grid), _lib.ptr(self._yee),
                                                   _lib.ptr(cc), _lib.stream_handle()))

    # -- time evolution (fields.py:568-647) -----------------------------------
    def evolve_electrodynamic_fields(self, J, dt):
        """J is [3][Q] FP64 (already summed over species)."""
        s = self._solver
        E_at_n = self._cc[:3].clone()
        self._cc_at_n[:3] = E_at_n
        B_at_n_minus_half = self._cc_at_n_plus_half[3:].clone()
        B_at_n_plus_half = self._cc[3:].clone()
        self._cc_at_n[3:] = 0.5 * (B_at_n_minus_half + B_at_n_plus_half)
        self._cc_at_n_plus_half[3:] = B_at_n_plus_half
        _lib.check(s._L.boltx_fdtd_step(ctypes.byref(s._grid), _lib.ptr(self._yee), _lib.ptr(J),
                                        float(dt), float(self.params.eps), float(self.params.mu),
                                        _lib.stream_handle()))
        self._yee_to_cc(self._cc)
        E_at_n_plus_one = self._cc[:3].clone()
        self._cc_at_n_plus_half[:3] = 0.5 * (E_at_n + E_at_n_plus_one)
        self.time_elapsed += dt

    def evolve_once_per_step(self, J, dt):
        """The once-per-step leapfrog of SSP-RK3 (J = J^{n+1/2}, [3][Q] FP64
        summed over species): from (E^n, B^{n+1/2}) to (E^{n+1}, B^{n+3/2})
        with the time levels of fields.py:568-647 in separate arrays:
        ``_cc_at_n`` = (E^{n+1}, (B^{n+1/2} + B^{n+3/2}) / 2) for stage 2 and the
        next step's stage 1, ``_cc_at_n_plus_half`` = ((E^n + E^{n+1}) / 2,
        B^{n+1/2}) for stage 3."""
        s = self._solver
        E_n = self._cc[:3].clone()
        B_n_plus_half = self._cc[3:].clone()
        _lib.check(s._L.boltx_fdtd_step(ctypes.byref(s._grid), _lib.ptr(self._yee), _lib.ptr(J),
                                        float(dt), float(self.params.eps), float(self.params.mu),
                                        _lib.stream_handle()))
        self._yee_to_cc(self._cc)
        self._cc_at_n = torch.cat([self._cc[:3], 0.5 * (B_n_plus_half + self._cc[3:])])
        self._cc_at_n_plus_half = torch.cat([0.5 * (E_n + self._cc[:3]), B_n_plus_half])
        self.time_elapsed += dt

    def _take_for_C_p(self):
        """The [6][Q] array one C_p evaluation sees: with the toggle coupling
        the time level at_n selects, then the toggle of fields.py:685-688;
        with the once-per-step coupling the stage's fields (the fields at n
        outside a step), no toggle."""
        if self.coupling == 'once-per-step':
            return self.stage_fields if self.stage_fields is not None else self._cc_at_n
        if not self._fdtd:
            src = self._cc
        else:
            src = self._cc_at_n if self.at_n else self._cc_at_n_plus_half
        if self.params.solver_method_in_p == 'FVM':
            self.at_n = not self.at_n
        return src

    def get_fields(self):
        """fields.py:649-690: cell-centred (E1, E2, E3, B1, B2, B3) shaped
        (1, 1, N_q1 + 2 N_g, N_q2 + 2 N_g), toggling the time level."""
        src = self._take_for_C_p()
        return tuple(self._solver._q_padded(src[k]) for k in range(6))

    # reference attribute names, as padded (6, 1, Q1g, Q2g) copies
    @property
    def cell_centered_EM_fields(self):
        return torch.stack([self._solver._q_padded(self._cc[k])[0] for k in range(6)])

    @property
    def yee_grid_EM_fields(self):
        return torch.stack([self._solver._q_padded(self._yee[k])[0] for k in range(6)])
