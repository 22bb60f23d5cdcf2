"""Electrodynamic fields for the FVM step: mirror of
bolt/lib/nonlinear/fields/fields.py:19-690 (FDTD branch) on the GPU.

Field arrays are FP64 [6][Q] device tensors (E1, E2, E3, B1, B2, B3; Q interior
q-cells, periodic, no ghost zones).  The staggered bookkeeping of
evolve_electrodynamic_fields (fields.py:568-647) is kept as in the reference,
including fields.py:493-494, which binds ``cell_centered_EM_fields_at_n`` and
``cell_centered_EM_fields_at_n_plus_half`` to the *same* array as
``cell_centered_EM_fields`` (component writes through any of the three names
change all of them).  The FDTD update and the Yee-to-centre averaging run in
the HIP library.
"""
import ctypes

import torch

from . import _lib

F64 = torch.float64


class fields_solver(object):
    """fields.py:19-690 for fields_type 'electrodynamic', fields_solver 'fdtd'."""

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
        if p.fields_initialize != 'user-defined':
            raise NotImplementedError(
                "fields_initialize = %r: only 'user-defined' is supported" % p.fields_initialize)
        self._initialize_electric_fields()
        self._initialize_magnetic_fields()
        self._yee_to_cc(cc)
        # fields.py:493-494: all three names refer to one object.
        self._cc = cc
        self._cc_at_n = cc
        self._cc_at_n_plus_half = cc

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

    def _take_for_C_p(self):
        """The [6][Q] array one C_p evaluation sees, then the toggle of
        fields.py:685-688."""
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
