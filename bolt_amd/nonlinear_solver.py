"""Drop-in mirror of bolt/lib/nonlinear/nonlinear_solver.py (FVM strategy),
with the time step running in the gfx950 HIP library.

User code keeps the reference interface: ``nonlinear_solver(system)``,
``strang_timestep(dt)`` (and lie/swss/jia, identical for FVM in q and p,
timestep.py:13-15), ``compute_moments(name, f=None)``, ``f``, ``time_elapsed``,
the grids ``q1_center`` ... ``p3_center`` and ``fields_solver``.  Arrays are
torch tensors on the GPU in the reference's axis order: f is
(N_p, N_s, N_q1 + 2 N_g, N_q2 + 2 N_g), q-grids are (1, 1, Q1g, Q2g) and
p-grids are (N_p, N_s, 1, 1).

Internally f is stored without ghost zones as a contiguous [N_q2][N_q1][N_p]
array (periodic q is handled by modular indexing in the kernels), FP64 or -- with
``params.precision = 'fp32-bulk'`` -- FP32 with FP64 moments.  Non-periodic q
boundaries ('mirror', 'dirichlet', 'mirror+dirichlet'; boundaries.py) keep
their ghost zones as extra q-rows after the interior, refilled on the device
before every RHS evaluation.  There is no CPU path: construction fails if the
HIP library or the GPU is missing.

Additions to the reference's parameters (all optional):
  ``riemann_solver_in_q/p``: 'upwind-flux' (reference) or 'lax-friedrichs';
  ``time_stepper_fvm``: 'RK2' (reference, fvm_operator.py:6-35) or 'SSP-RK3';
  ``precision``: 'fp64' (default) or 'fp32-bulk';
  ``memory_layout``: 'auto' (default), 'reference' or 'q1-lines'.
"""
import ctypes
import math
import types

import torch

from . import _lib
from . import file_io as _file_io
from .fields import fields_solver as _fields_solver
from .src.nonrelativistic_boltzmann import advection_terms as _adv
from .src.nonrelativistic_boltzmann import collision_operator as _coll
from .src.nonrelativistic_boltzmann import moments as _mom

F64 = torch.float64

_PLUGIN_PKG = 'nonrelativistic_boltzmann'


def _is_builtin(fn, builtin, module):
    """True for the built-in plugin function ``builtin`` itself, or for a
    function of the same name from a module of the same dotted name ending
    (``...nonrelativistic_boltzmann.<module>``): the reference's own plugin
    module (bolt/src/nonrelativistic_boltzmann) or a copy of it, which the
    fused kernels evaluate in-register."""
    if fn is builtin:
        return True
    mod = getattr(fn, '__module__', None) or ''
    suffix = _PLUGIN_PKG + '.' + module
    return getattr(fn, '__name__', None) == builtin.__name__ and \
        (mod == suffix or mod.endswith('.' + suffix))


def _is_builtin_module(m, builtin):
    name = getattr(m, '__name__', '') or ''
    suffix = _PLUGIN_PKG + '.' + builtin.__name__.rsplit('.', 1)[-1]
    return m is builtin or name == suffix or name.endswith('.' + suffix)


class _Species(object):
    """The state one species needs on the device: its grid (own p-range and
    spacing, calculate_p.py:33-57), physics (own mass, charge, collision
    time), f storage, moment block, workspace and work arrays.  Species are
    coupled only through the fields (J and rho summed over species,
    fields.py:562, 604-606)."""

    def __init__(self):
        self.grid = self.phys = self.f = self.mom = self.ws = None
        self.bufs = {}
        self.mass = self.charge = self.tau = 1.0
        self.tau_field = None  # (P, S, Q1g, Q2g) FP64 when tau varies over the grid
        self.dp = [1.0, 1.0, 1.0]


def _species_attr(name, doc):
    def get(self):
        return getattr(self._spc[self._cur], name)

    def put(self, value):
        setattr(self._spc[self._cur], name, value)
    return property(get, put, doc=doc)


class nonlinear_solver(object):
    """nonlinear_solver.py:55-521 for the FVM path on one GPU."""

    # the current species' state (species 0 unless a loop over species is
    # running); single-species code uses these names throughout
    _grid = _species_attr('grid', 'boltx_grid of the current species')
    _phys = _species_attr('phys', 'boltx_physics of the current species')
    _f = _species_attr('f', 'f of the current species in the storage layout')
    _mom = _species_attr('mom', 'FP64 moment block [12][Q] of the current species')
    _ws = _species_attr('ws', 'workspace of the current species')
    _bufs = _species_attr('bufs', 'work arrays of the current species')
    _mass = _species_attr('mass', 'mass of the current species')
    _charge = _species_attr('charge', 'charge of the current species')
    _tau = _species_attr('tau', 'collision time of the current species')
    _dp = _species_attr('dp', '[dp1, dp2, dp3] of the current species')

    def __init__(self, physical_system, performance_test_flag=False):
        self._L = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError('bolt_amd runs on a HIP GPU; none is visible (no CPU path)')
        self.device = torch.device('cuda', torch.cuda.current_device())
        ps = self.physical_system = physical_system
        up = ps.params
        # nonlinear_solver.py:100-121
        self.q1_start, self.q1_end = ps.q1_start, ps.q1_end
        self.q2_start, self.q2_end = ps.q2_start, ps.q2_end
        self.N_q1, self.dq1 = ps.N_q1, ps.dq1
        self.N_q2, self.dq2 = ps.N_q2, ps.dq2
        self.N_p1, self.N_p2, self.N_p3 = ps.N_p1, ps.N_p2, ps.N_p3
        self.N_ghost = ps.N_ghost
        self.boundary_conditions = bc = ps.boundary_conditions
        self.N_species = ps.N_species
        self.performance_test_flag = performance_test_flag
        # nonlinear_solver.py:177-249 and boundaries.py:415-577
        sides = ('in_q1_left', 'in_q1_right', 'in_q2_bottom', 'in_q2_top')
        for side in sides:
            if getattr(bc, side) == 'shearing-box':
                raise NotImplementedError('shearing-box boundaries are not provided')
            if getattr(bc, side) not in _lib.BC:
                raise NotImplementedError('Unavailable/Invalid boundary condition')
        for a, b in (('in_q1_left', 'in_q1_right'), ('in_q2_bottom', 'in_q2_top')):
            for kind in ('periodic', 'none'):
                if (getattr(bc, a) == kind) != (getattr(bc, b) == kind):
                    raise Exception('%s boundary conditions need to be applied to both the '
                                    'boundaries of a particular axis' % kind.capitalize())
        self.N_g1 = 0 if bc.in_q1_left == 'none' else self.N_ghost
        self.N_g2 = 0 if bc.in_q2_bottom == 'none' else self.N_ghost
        self._bc = [_lib.BC[getattr(bc, side)] for side in sides]
        self._ghosted = (self._bc[0] != 0, self._bc[2] != 0)
        if bc.in_q1_left == 'none' and self.N_q1 != 1:
            raise NotImplementedError("'none' boundaries are supported on singleton axes")
        if bc.in_q2_bottom == 'none' and self.N_q2 != 1:
            raise NotImplementedError("'none' boundaries are supported on singleton axes")

        # params as the callbacks see them: mass / charge as (1, N_s, 1, 1)
        # tensors (nonlinear_solver.py:133-145), the user's module untouched.
        self.params = types.SimpleNamespace(
            **{k: getattr(up, k) for k in dir(up) if not k.startswith('__')})
        S = self.N_species
        if len(up.mass) != S or len(up.charge) != S:
            raise Exception('params.mass and params.charge need one entry per species')
        # per-species state (grid, physics, storage, moments, workspace); the
        # solver's single-species attributes (_grid, _phys, _f, _mom, ...)
        # refer to species _cur
        self._spc = [_Species() for _ in range(S)]
        self._cur = 0
        for k, sp in enumerate(self._spc):
            sp.mass, sp.charge = float(up.mass[k]), float(up.charge[k])
        self.params.mass = torch.tensor([list(up.mass)], dtype=F64,
                                        device=self.device).reshape(1, -1, 1, 1)
        self.params.charge = torch.tensor([list(up.charge)], dtype=F64,
                                          device=self.device).reshape(1, -1, 1, 1)
        p = self.params

        self._configure(ps, p)
        self._build_grids(ps)
        if self._general and ((self.N_q1 > 1 and self.N_g1 < 1) or
                              (self.N_q2 > 1 and self.N_g2 < 1)):
            raise NotImplementedError('user-supplied terms need ghost zones on every q axis '
                                      'with more than one cell (their face values reach one '
                                      'cell past the interior)')
        for k, sp in enumerate(self._spc):
            sp.grid = _lib.Grid(self.N_q1, self.N_q2, self.N_p1, self.N_p2, self.N_p3, 1,
                                self._layout, float(self.dq1), float(self.dq2), sp.dp[0],
                                sp.dp[1], sp.dp[2], float(ps.p1_start[k]),
                                float(ps.p2_start[k]), float(ps.p3_start[k]),
                                int(self.N_ghost), (ctypes.c_int32 * 4)(*self._bc))
            sp.phys = _lib.Physics(self._recon_q, self._recon_p, self._riemann_q,
                                   self._riemann_p, int(self._source_on), int(self._p_adv),
                                   int(getattr(p, 'p_dim', 1)), self._precision,
                                   sp.mass, sp.charge,
                                   float(getattr(p, 'boltzmann_constant', 1.0)),
                                   float(sp.tau))
        Q, P = self.N_q1 * self.N_q2, self.N_p1 * self.N_p2 * self.N_p3
        self._Q, self._P = Q, P
        # ghost q-rows of the non-periodic axes after the Q interior rows
        self._G = (2 * self.N_ghost * self.N_q2 if self._ghosted[0] else 0) + \
                  (2 * self.N_ghost * self.N_q1 if self._ghosted[1] else 0)
        self._fshape = (self.N_q2, self.N_q1, P) if self._layout == 0 else \
            (self.N_q2, P, self.N_q1)
        self._dir_zero = None
        for sp in self._spc:
            sp.bufs = {}
            sp.mom = torch.zeros(_lib.MOM_ROWS, Q, dtype=F64, device=self.device)
            nws = int(self._L.boltx_workspace_size(ctypes.byref(sp.grid), ctypes.byref(sp.phys)))
            sp.ws = None
            if nws > 0:
                sp.ws = torch.empty(nws, dtype=torch.uint8, device=self.device)
                _lib.check(self._L.boltx_workspace_init(ctypes.byref(sp.grid),
                                                        ctypes.byref(sp.phys),
                                                        _lib.ptr(sp.ws), _lib.stream_handle()))
        self._flag = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.time_elapsed = 0
        self.dt = None
        self._A_q, self._C_q = ps.A_q, ps.C_q
        self._A_p, self._C_p = ps.A_p, ps.C_p
        self._source = ps.source
        self._initialize(p)
        self._timings = {'time_ts': 0.0, 'steps': 0}
        # list of (start, end) HIP events around every stage-kernel call, or None
        self._stage_events = None

    # -- configuration ------------------------------------------------------
    def _configure(self, ps, p):
        if getattr(p, 'solver_method_in_q', 'FVM') != 'FVM':
            raise NotImplementedError('only the FVM strategy is provided')
        rq = p.reconstruction_method_in_q
        rp = getattr(p, 'reconstruction_method_in_p', rq)
        if rq not in _lib.RECON or rp not in _lib.RECON:
            raise NotImplementedError('Reconstruction method invalid/not-implemented')
        self._recon_q, self._recon_p = _lib.RECON[rq], _lib.RECON[rp]
        rsq = getattr(p, 'riemann_solver_in_q', 'upwind-flux')
        rsp = getattr(p, 'riemann_solver_in_p', 'upwind-flux')
        if rsq not in _lib.RIEMANN or rsp not in _lib.RIEMANN:
            raise NotImplementedError('Riemann solver invalid/not-implemented')
        self._riemann_q, self._riemann_p = _lib.RIEMANN[rsq], _lib.RIEMANN[rsp]
        prec = getattr(p, 'precision', 'fp64')
        if prec not in _lib.PRECISION:
            raise NotImplementedError('precision must be fp64 or fp32-bulk')
        self._precision = _lib.PRECISION[prec]
        self._dtype = torch.float64 if prec == 'fp64' else torch.float32
        self._stepper = getattr(p, 'time_stepper_fvm', 'RK2')
        # memory layout of f: the q1-line layout [q2][p][q1] for fields-off
        # problems whose q1 extent fills whole 64-wide waves, else the
        # reference byte order [q2][q1][p]
        lay = getattr(p, 'memory_layout', 'auto')
        fields = bool(getattr(p, 'fields_enabled', False))
        ghosted = any(self._ghosted)
        self._plugin_flags(ps, p)
        if lay == 'auto':
            lay = 'q1-lines' if (not fields and not ghosted and ps.N_q1 in (64, 128, 256)
                                 and not self._general) else 'reference'
        if lay not in _lib.LAYOUT:
            raise NotImplementedError("memory_layout must be 'auto', 'reference' or 'q1-lines'")
        if lay == 'q1-lines' and self._general:
            raise NotImplementedError('user-supplied C_q, C_p, source or a varying tau need '
                                      'the reference layout')
        if lay == 'q1-lines' and (fields or ps.N_q1 > 256):
            raise NotImplementedError('the q1-line layout needs fields off and N_q1 <= 256')
        if lay == 'q1-lines' and ghosted:
            raise NotImplementedError('non-periodic q boundaries need the reference layout')
        if ghosted:
            if fields and getattr(p, 'fields_type', None) == 'electrodynamic':
                # the reference's FDTD near a non-periodic boundary reads Yee ghost
                # values that its circular shifts of the padded arrays evolve
                # (fdtd_explicit.py:9-177) and that apply_bcs_fields only partly
                # resets (fields/boundaries.py:419-567)
                raise NotImplementedError('non-periodic q boundaries are provided with the '
                                          'fields off or electrostatic')
            if getattr(p, 'p_space_grid', 'cartesian') != 'cartesian' and \
                    any(k in (1, 3) for k in self._bc):
                raise NotImplementedError('Unsupported coordinate system in p_space')
            need = {'piecewise-constant': 1, 'minmod': 2}.get(p.reconstruction_method_in_q, 3)
            if self.N_ghost < need:
                raise NotImplementedError('N_ghost = %d is below the reach of %s'
                                          % (self.N_ghost, p.reconstruction_method_in_q))
        self._layout = _lib.LAYOUT[lay]
        if self._stepper not in ('RK2', 'SSP-RK3'):
            raise NotImplementedError('time_stepper_fvm must be RK2 or SSP-RK3')
        self._fields_on = bool(getattr(p, 'fields_enabled', False))
        self._p_adv = self._fields_on and getattr(p, 'solver_method_in_p', 'FVM') == 'FVM'
        self._es = False
        if self._fields_on:
            ftype = getattr(p, 'fields_type', None)
            fsolver = getattr(p, 'fields_solver', '').lower()
            if ftype == 'electrodynamic':
                if fsolver != 'fdtd':
                    raise NotImplementedError('electrodynamic fields need the FDTD solver')
            elif ftype == 'electrostatic':
                if fsolver != 'fft':
                    raise NotImplementedError('electrostatic fields need the FFT solver')
                self._es = True
            else:
                raise NotImplementedError('fields_type %r is not provided' % ftype)
            if getattr(p, 'hybrid_model_enabled', False):
                raise NotImplementedError('the hybrid model is not provided')
            if self._user_cq and ftype == 'electrodynamic':
                raise NotImplementedError('the currents of the FDTD update take their face '
                                          'states from the built-in C_q')

    def _plugin_flags(self, ps, p):
        """Which terms are the built-in plugins (bolt_amd.src or the
        reference's bolt.src.nonrelativistic_boltzmann modules: evaluated
        in-register by the fused kernels) and which are the user's callbacks
        (evaluated with torch on the device each RHS and fed to the generic
        stage kernel, df_dt_fvm.py:67-97, 251-267).  A collision time that
        varies over the grid also takes the generic kernel."""
        self._source_on = bool(getattr(p, 'source_enabled', False))
        fields = bool(getattr(p, 'fields_enabled', False))
        self._user_cq = not _is_builtin(ps.C_q, _adv.C_q, 'advection_terms')
        self._user_cp = fields and getattr(p, 'solver_method_in_p', 'FVM') == 'FVM' and \
            not _is_builtin(ps.C_p, _adv.C_p, 'advection_terms')
        self._user_src = self._source_on and \
            not _is_builtin(ps.source, _coll.BGK, 'collision_operator')
        self._builtin_moments = _is_builtin_module(ps.moments, _mom)
        if self._source_on and not self._user_src and not self._builtin_moments:
            raise NotImplementedError('the built-in BGK source needs the built-in moments module')
        self._tau_varies = False  # set by _build_grids
        self._general = self._user_cq or self._user_cp or self._user_src

    def _build_grids(self, ps):
        """calculate_q.py:7-41 and calculate_p.py:8-57 as device tensors."""
        dev = self.device
        i1 = torch.arange(-self.N_g1, self.N_q1 + self.N_g1, dtype=F64, device=dev)
        i2 = torch.arange(-self.N_g2, self.N_q2 + self.N_g2, dtype=F64, device=dev)
        Q1g, Q2g = len(i1), len(i2)
        q1lb = (self.q1_start + i1 * self.dq1).reshape(1, 1, Q1g, 1).expand(1, 1, Q1g, Q2g)
        q2lb = (self.q2_start + i2 * self.dq2).reshape(1, 1, 1, Q2g).expand(1, 1, Q1g, Q2g)
        self.q1_left_bot, self.q2_left_bot = q1lb.contiguous(), q2lb.contiguous()
        self.q1_center_bot = self.q1_left_bot + 0.5 * self.dq1
        self.q2_center_bot = self.q2_left_bot
        self.q1_left_center = self.q1_left_bot
        self.q2_left_center = self.q2_center_bot  # calculate_q.py:37
        self.q1_center = self.q1_left_bot + 0.5 * self.dq1
        self.q2_center = self.q2_left_bot + 0.5 * self.dq2
        S = self.N_species
        N = [self.N_p1, self.N_p2, self.N_p3]
        meshes = {0.5: [], 0.0: []}
        for k, sp in enumerate(self._spc):
            sp.dp = [float(ps.dp1[k]), float(ps.dp2[k]), float(ps.dp3[k])]
            starts = [float(ps.p1_start[k]), float(ps.p2_start[k]), float(ps.p3_start[k])]
            for off in (0.5, 0.0):
                ax = [starts[a] + (off + torch.arange(N[a], dtype=F64, device=dev)) * sp.dp[a]
                      for a in range(3)]
                g3, g2, g1 = torch.meshgrid(ax[2], ax[1], ax[0], indexing='ij')
                meshes[off].append([g.reshape(-1) for g in (g1, g2, g3)])

        def mesh(off):
            # (P, S, 1, 1) per axis: species k in column k (calculate_p.py:33-57)
            return [torch.stack([m[a] for m in meshes[off]], dim=1).reshape(-1, S, 1, 1)
                    .contiguous() for a in range(3)]

        self.p1_center, self.p2_center, self.p3_center = mesh(0.5)
        self.p1_left, self.p2_bottom, self.p3_back = mesh(0.0)
        self.p2_left, self.p3_left = self.p2_center, self.p3_center
        self.p1_bottom, self.p3_bottom = self.p1_center, self.p3_center
        self.p1_back, self.p2_back = self.p1_center, self.p2_center
        self.dp1 = torch.tensor([sp.dp[0] for sp in self._spc], dtype=F64,
                                device=dev).reshape(1, S, 1, 1)
        self.dp2 = torch.tensor([sp.dp[1] for sp in self._spc], dtype=F64,
                                device=dev).reshape(1, S, 1, 1)
        self.dp3 = torch.tensor([sp.dp[2] for sp in self._spc], dtype=F64,
                                device=dev).reshape(1, S, 1, 1)
        # collision time of the built-in BGK: params.tau(q1, q2, p1, p2, p3)
        # (collision_operator.py:119), which has no time argument, evaluated
        # once; uniform per species -> a scalar for the fused kernels, else a
        # per-cell field for the generic kernel
        for sp in self._spc:
            sp.tau, sp.tau_field = 1.0, None
        if self._source_on and not self._user_src:
            tau = self.params.tau(self.q1_center, self.q2_center, self.p1_center,
                                  self.p2_center, self.p3_center)
            tau = torch.as_tensor(tau, dtype=F64, device=dev)
            tau = torch.broadcast_to(tau, (self.p1_center.shape[0], S) +
                                     tuple(self.q1_center.shape[2:]))
            for k, sp in enumerate(self._spc):
                t0 = float(tau[:, k].reshape(-1)[0])
                sp.tau = t0
                if not bool((tau[:, k] == t0).all()):
                    sp.tau_field = tau
                    self._tau_varies = True
            if self._tau_varies:
                self._general = True
                if self._layout == _lib.LAYOUT['q1-lines']:
                    if getattr(self.params, 'memory_layout', 'auto') == 'q1-lines':
                        raise NotImplementedError('a collision time that varies over the grid '
                                                  'needs the reference layout')
                    self._layout = _lib.LAYOUT['reference']

    # -- layout helpers -------------------------------------------------------
    def _q_interior(self, x):
        """(..., Q1g, Q2g) -> interior (..., N_q1, N_q2)."""
        g1, g2 = self.N_g1, self.N_g2
        return x[..., g1:g1 + self.N_q1, g2:g2 + self.N_q2]

    def _q_padded(self, flat):
        """[Q] (index i1 + N_q1 i2) -> (1, 1, Q1g, Q2g) with periodic ghosts."""
        a = flat.reshape(self.N_q2, self.N_q1).t()
        i1 = (torch.arange(-self.N_g1, self.N_q1 + self.N_g1, device=a.device)) % self.N_q1
        i2 = (torch.arange(-self.N_g2, self.N_q2 + self.N_g2, device=a.device)) % self.N_q2
        return a[i1][:, i2].reshape(1, 1, len(i1), len(i2))

    def _alloc_f(self, dtype=None):
        """An f-sized array in the storage layout whose memory continues with
        the G ghost q-rows of the non-periodic axes (the view covers the
        interior; the kernels address the rows after it)."""
        dtype = dtype or self._dtype
        n = self._Q * self._P
        store = torch.empty((self._Q + self._G) * self._P, dtype=dtype, device=self.device)
        return store[:n].view(self._fshape)

    def _store(self, x):
        """x (storage layout) as an array with room for the ghost rows."""
        if self._G == 0:
            return x.contiguous().to(self._dtype)
        out = self._alloc_f()
        out.copy_(x)
        return out

    def _to_internal(self, f_ref):
        """(P, S, Q1g, Q2g) -> the contiguous storage layout ([N_q2][N_q1][P]
        or [N_q2][P][N_q1]) in the storage dtype."""
        f_ref = torch.as_tensor(f_ref, device=self.device)
        f_ref = torch.broadcast_to(f_ref, (self._P, 1) + tuple(self.q1_center.shape[2:]))
        inner = self._q_interior(f_ref)[:, 0]  # (P, N_q1, N_q2)
        perm = (2, 1, 0) if self._layout == 0 else (2, 0, 1)
        return self._store(inner.permute(*perm))

    def _padded_rows(self):
        """Row of the [Q + G] storage behind every padded (Q1g, Q2g) cell: the
        periodic image, a ghost row, or -1 for the corners (no stencil reads
        them; they read as zero)."""
        if getattr(self, '_rows_cache', None) is not None:
            return self._rows_cache
        n1, n2, g = self.N_q1, self.N_q2, self.N_ghost
        i1 = torch.arange(-self.N_g1, n1 + self.N_g1).reshape(-1, 1)
        i2 = torch.arange(-self.N_g2, n2 + self.N_g2).reshape(1, -1)
        in1 = (i1 >= 0) & (i1 < n1)
        in2 = (i2 >= 0) & (i2 < n2)
        j1, j2 = i1 % n1, i2 % n2
        rows = j1 + n1 * j2
        Q = self._Q
        # a periodic axis's ghost cells are its periodic images, also beside
        # the other axis's ghost rows
        if self._ghosted[0]:
            c = torch.where(i1 < 0, i1 + g, i1 - n1)
            r = Q + torch.where(i1 < 0, 0, g * n2) + c * n2 + j2
            rows = torch.where(~in1 & (in2 | (not self._ghosted[1])), r, rows)
        if self._ghosted[1]:
            c = torch.where(i2 < 0, i2 + g, i2 - n2)
            off = Q + (2 * g * n2 if self._ghosted[0] else 0)
            r = off + torch.where(i2 < 0, 0, g * n1) + c * n1 + j1
            rows = torch.where((in1 | (not self._ghosted[0])) & ~in2, r, rows)
        if any(self._ghosted):
            outside = (~in1 if self._ghosted[0] else torch.zeros_like(in1)) & \
                      (~in2 if self._ghosted[1] else torch.zeros_like(in2))
            rows = torch.where(outside, -1, rows)
        self._rows_cache = rows.to(self.device)
        return self._rows_cache

    def _to_reference(self, f_int):
        """Storage layout -> (P, 1, Q1g, Q2g): periodic ghost zones from the
        interior, non-periodic ones from the ghost rows (as last filled)."""
        if self._G == 0:
            i1 = torch.arange(-self.N_g1, self.N_q1 + self.N_g1, device=f_int.device) % self.N_q1
            i2 = torch.arange(-self.N_g2, self.N_q2 + self.N_g2, device=f_int.device) % self.N_q2
            if self._layout == 0:
                return f_int[i2][:, i1].permute(2, 1, 0).unsqueeze(1)
            return f_int[i2][:, :, i1].permute(1, 2, 0).unsqueeze(1)
        n = (self._Q + self._G) * self._P
        full = torch.as_strided(f_int, (self._Q + self._G, self._P), (self._P, 1))
        assert full.untyped_storage().nbytes() >= n * f_int.element_size()
        rows = self._padded_rows()
        out = full[rows.clamp(min=0)]  # (Q1g, Q2g, P)
        out = torch.where((rows >= 0).unsqueeze(-1), out, torch.zeros((), dtype=out.dtype,
                                                                      device=out.device))
        return out.permute(2, 0, 1).unsqueeze(1)

    def _fill_ghosts(self, u):
        """communicate_f + apply_bcs_f (communicate.py:8-38, boundaries.py:
        415-577) on the ghost rows of u before a RHS evaluation.  Dirichlet
        sides call the user's f_left / f_right / f_bottom / f_top(f, t, q1,
        q2, p1, p2, p3, params) with f in the reference layout, its ghost
        zones zero or mirrored (the user's values are not in yet), and take
        their values where the characteristic enters."""
        if self._G == 0:
            return
        dirs = (ctypes.c_void_p * 4)()
        keep = []
        dsides = [k for k in range(4) if self._bc[k] in (2, 3)]
        if dsides:
            if self._dir_zero is None:
                n = self.N_ghost * max(self.N_q1, self.N_q2) * self._P
                self._dir_zero = torch.zeros(n, dtype=self._dtype, device=self.device)
            for k in dsides:
                dirs[k] = self._dir_zero.data_ptr()
            _lib.check(self._L.boltx_fill_ghosts(ctypes.byref(self._grid), ctypes.byref(self._phys),
                                                 _lib.ptr(u), dirs, _lib.stream_handle()))
            fpad = self._to_reference(u).to(F64)
            bcm = self.boundary_conditions
            args = (fpad, self.time_elapsed, self.q1_center, self.q2_center, self.p1_center,
                    self.p2_center, self.p3_center, self.params)
            g1, g2, g = self.N_g1, self.N_g2, self.N_ghost
            full = (self._P, 1) + tuple(self.q1_center.shape[2:])
            for k in dsides:
                fn = getattr(bcm, ('f_left', 'f_right', 'f_bottom', 'f_top')[k])
                val = torch.broadcast_to(torch.as_tensor(fn(*args), dtype=F64,
                                                         device=self.device), full)[:, 0]
                if k == 0:
                    blk = val[:, :g, g2:g2 + self.N_q2]                  # (P, c, i2)
                elif k == 1:
                    blk = val[:, g1 + self.N_q1:g1 + self.N_q1 + g, g2:g2 + self.N_q2]
                elif k == 2:
                    blk = val[:, g1:g1 + self.N_q1, :g].transpose(1, 2)  # (P, c, i1)
                else:
                    blk = val[:, g1:g1 + self.N_q1, g2 + self.N_q2:g2 + self.N_q2 + g].transpose(1, 2)
                blk = blk.permute(1, 2, 0).contiguous().to(self._dtype)  # [c][i_other][P]
                keep.append(blk)
                dirs[k] = blk.data_ptr()
        _lib.check(self._L.boltx_fill_ghosts(ctypes.byref(self._grid), ctypes.byref(self._phys),
                                             _lib.ptr(u), dirs, _lib.stream_handle()))

    def _species_loop(self):
        """Iterate over the species, making each the current one; species 0
        is current again afterwards."""
        try:
            for k in range(len(self._spc)):
                self._cur = k
                yield k
        finally:
            self._cur = 0

    @property
    def f(self):
        """The distribution function in the reference's (P, S, Q1g, Q2g)
        layout (a copy): periodic ghost zones filled from the interior,
        non-periodic ones as the boundary conditions fill them."""
        parts = []
        for _ in self._species_loop():
            self._fill_ghosts(self._f)
            parts.append(self._to_reference(self._f))
        return parts[0] if len(parts) == 1 else torch.cat(parts, dim=1)

    @f.setter
    def f(self, value):
        value = torch.as_tensor(value, device=self.device)
        value = torch.broadcast_to(value, (self._P, self.N_species) +
                                   tuple(self.q1_center.shape[2:]))
        for k in self._species_loop():
            self._f = self._to_internal(value[:, k:k + 1])
            self._update_moments()

    # -- initialisation (nonlinear_solver.py:464-500) -------------------------
    def _initialize(self, params):
        f = self.physical_system.initial_conditions.initialize_f(
            self.q1_center, self.q2_center, self.p1_center, self.p2_center,
            self.p3_center, params)
        self.f = torch.as_tensor(f, dtype=F64, device=self.device)
        self.f_initial = self._f.clone()
        self.fields_solver = _fields_solver(self) if self._fields_on else None

    def _update_moments(self):
        """The moment block of the current species' f."""
        _lib.check(self._L.boltx_moments(ctypes.byref(self._grid), ctypes.byref(self._phys),
                                         _lib.ptr(self._f), _lib.ptr(self._mom),
                                         _lib.ptr(self._ws), _lib.stream_handle()))

    def _rho(self, moms=None):
        """Charge density summed over species (fields.py:562), [Q] FP64, from
        the species' moment blocks (or ``moms``, one per species)."""
        moms = moms or [sp.mom for sp in self._spc]
        rho = self._spc[0].charge * moms[0][0]
        for sp, m in zip(self._spc[1:], moms[1:]):
            rho = rho + sp.charge * m[0]
        return rho

    def _buf(self, name, dtype=None, shape=None):
        """A cached work array; f-sized ones (no shape given) have room for
        the ghost rows."""
        dtype = dtype or self._dtype
        fsized = shape is None
        shape = shape or tuple(self._f.shape)
        t = self._bufs.get(name)
        if t is None or t.dtype != dtype or tuple(t.shape) != tuple(shape):
            t = self._alloc_f(dtype) if fsized else torch.empty(shape, dtype=dtype,
                                                                 device=self.device)
            self._bufs[name] = t
        return t

    # -- the step -------------------------------------------------------------
    def _currents(self, us, name):
        """J ([3][Q] FP64) of the species' states ``us`` summed over species
        (df_dt_fvm.py:214-224, fields.py:604-606), into work array ``name``."""
        J = None
        for k in self._species_loop():
            Jk = self._buf('J', F64, (3, self._Q))
            _lib.check(self._L.boltx_currents(ctypes.byref(self._grid),
                                              ctypes.byref(self._phys), _lib.ptr(us[k]),
                                              _lib.ptr(Jk), _lib.stream_handle()))
            if J is None:
                J = self._spc[0].bufs.setdefault(name, torch.empty_like(Jk))
                J.copy_(Jk)
            else:
                J += Jk
        return J

    def _fields_for_stage(self, us, stage=None):
        """The field work of one RHS evaluation (df_dt_fvm.py:102-104,
        214-267), with J and rho summed over species (fields.py:562,
        604-606), and the three field arrays C_p1..C_p3 see.  ``us`` are the
        species' stage inputs; their moment blocks are ``_stage_moms`` (set by
        the caller).  Electrostatic: the solve from rho of the stage input
        (df_dt_fvm.py:235-242).  Electrodynamic with the reference's toggle
        (RK2): the FDTD update from J when at_n is False, then at_n toggles
        per C_p.  Electrodynamic under SSP-RK3 (``stage`` 0, 1, 2): stage 1
        sees the fields at n and keeps J(u^n); stage 2 takes J(u1), advances
        the fields once with their average J^{n+1/2} and sees the fields at
        n+1; stage 3 sees those at n+1/2."""
        F = [None, None, None]
        if not self._p_adv:
            return F
        fs = self.fields_solver
        if self._es:  # df_dt_fvm.py:235-242: rho of the stage input
            fs.compute_electrostatic_fields(self._rho(self._stage_moms))
        elif fs.coupling == 'once-per-step':
            if stage == 0:
                # J(u^n) comes out of this stage's kernels (_advance sums it)
                fs.stage_fields = fs._cc_at_n
            elif stage == 1:
                J_n = self._spc[0].bufs['J_n']
                J_1 = self._currents(us, 'J_1')
                fs.evolve_once_per_step(0.5 * (J_n + J_1), self.dt)
                fs.stage_fields = fs._cc_at_n
            elif stage == 2:
                fs.stage_fields = fs._cc_at_n_plus_half
        elif fs.at_n is False:
            fs.evolve_electrodynamic_fields(self._currents(us, 'J_sum'), self.dt)
        if self._user_cp:
            # df_dt_fvm.py:251-267: the user's C_p at the p-faces, each call
            # taking the fields through fields_solver.get_fields()
            t, p = self.time_elapsed, self.params
            self._cp_user = [
                self._C_p(t, self.q1_center, self.q2_center, self.p1_left, self.p2_left,
                          self.p3_left, fs, p)[0],
                self._C_p(t, self.q1_center, self.q2_center, self.p1_bottom, self.p2_bottom,
                          self.p3_bottom, fs, p)[1],
                self._C_p(t, self.q1_center, self.q2_center, self.p1_back, self.p2_back,
                          self.p3_back, fs, p)[2]]
            return F
        return [fs._take_for_C_p() for _ in range(3)]  # df_dt_fvm.py:251-267

    # -- user callbacks (generic stage kernel) ---------------------------------
    def _coeff(self, x, k, keep):
        """boltx_coeff of species k of a callback result ``x`` (broadcastable
        to the reference's (P, S, Q1g, Q2g)), with its origin at the first
        interior q-cell; broadcast axes keep a zero stride.  The FP64 view is
        appended to ``keep`` so it outlives the launch."""
        x = torch.as_tensor(x, dtype=F64, device=self.device)
        full = (self._P, self.N_species) + tuple(self.q1_center.shape[2:])
        v = torch.broadcast_to(x, full)[:, k]
        keep.append(v)
        off = v.stride(1) * self.N_g1 + v.stride(2) * self.N_g2
        return _lib.Coeff(v.data_ptr() + v.element_size() * off, v.stride(0), v.stride(1),
                          v.stride(2))

    def _callback_terms(self, us):
        """The user's terms of one RHS evaluation with stage inputs ``us``
        (one per species), as one boltx_coeffs per species: C_q at the left
        and bottom q-faces (df_dt_fvm.py:67-77), the source on the stage input
        (:91-97), the C_p of this RHS (from _fields_for_stage) and a varying
        collision time.  Also returns the arrays to keep alive."""
        t, p = self.time_elapsed, self.params
        keep = []
        per = [_lib.Coeffs() for _ in self._spc]
        if self._user_cq:
            C1 = self._C_q(t, self.q1_left_center, self.q2_left_center, self.p1_center,
                           self.p2_center, self.p3_center, p)[0]
            C2 = self._C_q(t, self.q1_center_bot, self.q2_center_bot, self.p1_center,
                           self.p2_center, self.p3_center, p)[1]
            for k, c in enumerate(per):
                c.c_q1, c.c_q2 = self._coeff(C1, k, keep), self._coeff(C2, k, keep)
        if self._user_src:
            parts = []
            for k in self._species_loop():
                self._fill_ghosts(us[k])
                parts.append(self._to_reference(us[k]))
            f_ref = torch.cat(parts, dim=1).to(F64)
            S = self._source(f_ref, t, self.q1_center, self.q2_center, self.p1_center,
                             self.p2_center, self.p3_center, self.compute_moments, p, False)
            for k, c in enumerate(per):
                c.source = self._coeff(S, k, keep)
        if self._user_cp:
            for k, c in enumerate(per):
                c.c_p1, c.c_p2, c.c_p3 = (self._coeff(x, k, keep) for x in self._cp_user)
        for k, (c, sp) in enumerate(zip(per, self._spc)):
            if sp.tau_field is not None:
                c.tau = self._coeff(sp.tau_field, k, keep)
        return per, keep

    def _stage_kernel(self, u, w, un, out, a, b, cdt, mom_in, mom_out, F, co=None, jout=None):
        """The current species' stage launch, out = a*un + b*(w + cdt*L(u)),
        with its ghost refresh (fvm_operator.py:22-33) and the moments of
        out; ``co`` (boltx_coeffs) carries the user's terms to the generic
        kernel; ``jout`` ([3][Q] FP64) receives the currents of u."""
        self._fill_ghosts(u)
        if co is not None:
            st = _lib.Stage(_lib.ptr(u), _lib.ptr(w), _lib.ptr(un), _lib.ptr(out),
                            float(a), float(b), float(cdt), _lib.ptr(mom_in), _lib.ptr(mom_out),
                            _lib.ptr(F[0]), _lib.ptr(F[1]), _lib.ptr(F[2]), None,
                            _lib.ptr(jout))
            _lib.check(self._L.boltx_fvm_stage_coeffs(
                ctypes.byref(self._grid), ctypes.byref(self._phys), ctypes.byref(st),
                ctypes.byref(co), _lib.stream_handle()))
            return
        # In the q1-line layout the moments of the output are a separate
        # streaming pass; issue it as its own call.
        split = self._layout == _lib.LAYOUT['q1-lines'] and mom_out is not None
        st = _lib.Stage(_lib.ptr(u), _lib.ptr(w), _lib.ptr(un), _lib.ptr(out),
                        float(a), float(b), float(cdt), _lib.ptr(mom_in),
                        None if split else _lib.ptr(mom_out),
                        _lib.ptr(F[0]), _lib.ptr(F[1]), _lib.ptr(F[2]), _lib.ptr(self._ws),
                        _lib.ptr(jout))
        ev = self._stage_events
        if ev is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e0.record()
        _lib.check(self._L.boltx_fvm_stage(ctypes.byref(self._grid), ctypes.byref(self._phys),
                                           ctypes.byref(st), _lib.stream_handle()))
        if ev is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            ev.append((e0, e1))
        if split:
            _lib.check(self._L.boltx_moments(ctypes.byref(self._grid), ctypes.byref(self._phys),
                                             _lib.ptr(out), _lib.ptr(mom_out), _lib.ptr(self._ws),
                                             _lib.stream_handle()))

    def _stage(self, u, w, un, out, a, b, cdt, mom_in, mom_out):
        """One RK stage of a one-species problem, out = a*un + b*(w + cdt*L(u))
        (df_dt_fvm.py:33-339 inside fvm_operator.py:22-33)."""
        self._stage_moms = [mom_in]
        F = self._fields_for_stage([u])
        co = self._callback_terms([u])[0][0] if self._general else None
        self._stage_kernel(u, w, un, out, a, b, cdt, mom_in, mom_out, F, co)

    def _advance(self, dt):
        """Advance f by dt without the host-side divergence check (the device
        flag is still set; see strang_timestep).  Each stage evaluates the
        fields once from every species, then runs every species' stage."""
        self.dt = dt
        if self._stepper == 'RK2':
            # fvm_operator.py:25-33
            plan = [('f', 'f', None, 'u1', 0.0, 1.0, dt / 2, 'mom', 'm1'),
                    ('u1', 'f', None, 'f', 0.0, 1.0, dt, 'm1', 'mom')]
        else:
            plan = [('f', 'f', None, 'u1', 0.0, 1.0, dt, 'mom', 'm1'),
                    ('u1', 'u1', 'f', 'u2', 0.75, 0.25, dt, 'm1', 'm2'),
                    ('u2', 'u2', 'f', 'f', 1 / 3, 2 / 3, dt, 'm2', 'mom')]

        def arr(name):
            if name is None:
                return None
            if name == 'f':
                return self._f
            if name == 'mom':
                return self._mom
            if name in ('m1', 'm2'):
                return self._buf(name, F64, (_lib.MOM_ROWS, self._Q))
            return self._buf(name)

        for stage, (u, w, un, out, a, b, cdt, mi, mo) in enumerate(plan):
            args = []
            for _ in self._species_loop():
                args.append([arr(x) for x in (u, w, un, out, mi, mo)])
            self._stage_moms = [x[4] for x in args]
            F = self._fields_for_stage([x[0] for x in args], stage)
            co = [None] * len(self._spc)
            if self._general:
                co, keep = self._callback_terms([x[0] for x in args])
            # SSP-RK3 with FDTD: stage 1's kernels also give J(u^n) (df_dt_fvm.py:214-224)
            fuse_j = stage == 0 and self.fields_solver is not None and \
                self.fields_solver.coupling == 'once-per-step'
            for k in self._species_loop():
                uu, ww, nn, oo, mmi, mmo = args[k]
                jk = self._buf('J_stage', F64, (3, self._Q)) if fuse_j else None
                self._stage_kernel(uu, ww, nn, oo, a, b, cdt, mmi, mmo, F, co[k], jk)
            if fuse_j:
                J = self._spc[0].bufs.setdefault('J_n', torch.empty_like(jk))
                J.copy_(self._spc[0].bufs['J_stage'])
                for sp in self._spc[1:]:
                    J += sp.bufs['J_stage']
        for _ in self._species_loop():
            _lib.check(self._L.boltx_check_finite(ctypes.byref(self._grid), _lib.ptr(self._mom),
                                                  _lib.ptr(self._flag), _lib.stream_handle()))
        if self.fields_solver is not None:
            self.fields_solver.stage_fields = None
        self.time_elapsed += dt

    def capture(self, n_steps, dt):
        """Record ``n_steps`` steps of ``dt`` (the launches strang_timestep
        issues) into a HIP graph (torch.cuda.CUDAGraph); each replay() then
        advances f by those steps with one graph launch.  Periodic q, fields
        off and the built-in terms only: the host-side work of the other
        cases (field time levels, user callbacks, boundary functions of the
        time) would not repeat on replay.  The launches use fixed buffers and
        no memsets, so every replay repeats the eager steps exactly."""
        if self._p_adv or self._general or self._G:
            raise NotImplementedError('graph capture takes periodic q, fields off and the '
                                      'built-in advection, source and collision time')
        # the stages' work arrays exist before the capture (arrays allocated
        # inside it would come from the graph's private pool)
        names = ('u1', 'm1') if self._stepper == 'RK2' else ('u1', 'u2', 'm1', 'm2')
        for _ in self._species_loop():
            for nm in names:
                if nm[0] == 'm':
                    self._buf(nm, F64, (_lib.MOM_ROWS, self._Q))
                else:
                    self._buf(nm)
        events, self._stage_events = self._stage_events, None
        torch.cuda.synchronize()
        t, d = self.time_elapsed, getattr(self, 'dt', None)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(int(n_steps)):
                self._advance(dt)
        self.time_elapsed, self.dt, self._stage_events = t, d, events
        self._graph, self._graph_steps, self._graph_dt = g, int(n_steps), float(dt)
        return g

    def replay(self):
        """Run the captured steps (see capture), without a host sync."""
        self._graph.replay()
        for _ in range(self._graph_steps):
            self.time_elapsed += self._graph_dt
        self.dt = self._graph_dt

    def check_divergence(self):
        """timestep.py:28-38 (reads the device flag: one host sync)."""
        if int(self._flag.item()) != 0:
            raise SystemExit('Solver Diverging!')

    def strang_timestep(self, dt):
        """timestep.py:101-160: FVM in q and p needs no splitting."""
        if self.performance_test_flag:
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
        self._advance(dt)
        self.check_divergence()
        if self.performance_test_flag:
            t1.record()
            t1.synchronize()
            self._timings['time_ts'] += t0.elapsed_time(t1) / 1e3
            self._timings['steps'] += 1

    lie_timestep = strang_timestep
    swss_timestep = strang_timestep
    jia_timestep = strang_timestep

    # -- moments (compute_moments.py:7-59) -----------------------------------
    def compute_moments(self, moment_name, f=None):
        """Moments shaped (1, N_s, Q1g, Q2g).  The built-in five come from the
        HIP reduction; other moment functions of the user's module are
        evaluated with torch on the device."""
        if f is None and moment_name in _mom._BUILTIN and self._builtin_moments:
            k = _mom._BUILTIN.index(moment_name)
            return torch.cat([self._q_padded(sp.mom[k]) for sp in self._spc], dim=1)
        fn = getattr(self.physical_system.moments, moment_name)
        ff = self.f if f is None else torch.as_tensor(f, device=self.device)
        return fn(ff.to(F64), self.p1_center, self.p2_center, self.p3_center,
                  self.dp3 * self.dp2 * self.dp1)

    # -- checkpoints (nonlinear_solver.py:513-519, file_io/dump.py, load.py) ---
    def dump_distribution_function(self, file_name):
        _file_io.dump_distribution_function(self, file_name)

    def dump_moments(self, file_name):
        _file_io.dump_moments(self, file_name)

    def dump_EM_fields(self, file_name):
        _file_io.dump_EM_fields(self, file_name)

    def dump_aux_arrays(self, arrays, name, file_name):
        _file_io.dump_aux_arrays(self, arrays, name, file_name)

    def load_distribution_function(self, file_name):
        _file_io.load_distribution_function(self, file_name)

    def load_EM_fields(self, file_name):
        _file_io.load_EM_fields(self, file_name)

    def print_performance_timings(self, N_iters=None):
        """performance_timings.py:9-165 (reduced: one GPU, one table row)."""
        n = self._timings['steps']
        if n == 0:
            print('no timed steps (construct with performance_test_flag=True)')
            return
        t = self._timings['time_ts'] / n
        cells = self._Q * self._P * self.N_species
        print('Time per step           : %.6e s' % t)
        print('Phase-space cell-updates/s: %.6e' % (cells / t))
        print('Spatial Zone Cycles/s   : %.6e' % (self._Q / t))
