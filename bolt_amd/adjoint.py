"""Gradients through the FVM time step: a hand-written adjoint of the whole RK
step behind ``torch.autograd.Function``, with stage checkpointing.

The reference has no adjoint.  ``evolve(solver, f0, tau, n_steps, dt)`` runs
``n_steps`` steps of the solver's own scheme (reconstruction, flux, BGK,
RK2 midpoint or SSP-RK3: fvm_operator.py:6-50, df_dt_fvm.py:33-97,
collision_operator.py:70-128) from the interior state ``f0`` with the
collision time ``tau`` and returns the final interior state; backpropagating
through it yields d/d f0 and d/d tau.

Memory: the forward pass keeps one FP64 state every ``checkpoint_every``
steps; the backward pass re-runs each segment from its checkpoint, keeping
that segment's stage inputs and moment blocks, and applies the stage adjoints
(``boltx_adjoint_stage``) in reverse.

Interior states are float64 tensors shaped ``(P, N_q1, N_q2)``: the
reference's ``(P, S, Q1g, Q2g)`` array without ghost zones and with the one
species axis dropped.  FP64 only; fields off; periodic q.
"""
import ctypes

import torch

from . import _lib

F64 = torch.float64


def _check_solver(s):
    if s._precision != _lib.PRECISION['fp64']:
        raise NotImplementedError("the adjoint runs in plain FP64 (precision='fp64')")
    if s._fields_on:
        raise NotImplementedError('the adjoint needs fields off')
    if s._G:
        raise NotImplementedError('the adjoint needs periodic q boundaries')
    if s.N_species != 1:
        raise NotImplementedError('the adjoint is provided for one species')
    if s._general:
        raise NotImplementedError('the adjoint differentiates the built-in C_q and BGK with a '
                                  'uniform tau')


def interior_to_internal(s, x):
    """(P, N_q1, N_q2) -> the solver's storage layout."""
    perm = (2, 1, 0) if s._layout == 0 else (2, 0, 1)
    return x.to(F64).permute(*perm).contiguous()


def internal_to_interior(s, y):
    """The solver's storage layout -> (P, N_q1, N_q2)."""
    perm = (2, 1, 0) if s._layout == 0 else (1, 2, 0)
    return y.permute(*perm).contiguous()


def _set_tau(s, tau):
    s._tau = float(tau)
    s._phys.tau = float(tau)


def _moments_of(s, f):
    m = torch.empty(_lib.MOM_ROWS, s._Q, dtype=F64, device=s.device)
    _lib.check(s._L.boltx_moments(ctypes.byref(s._grid), ctypes.byref(s._phys), _lib.ptr(f),
                                  _lib.ptr(m), _lib.ptr(s._ws), _lib.stream_handle()))
    return m


def _step_recorded(s, f, dt, need_out=True):
    """One step from f (not modified): returns (f_next, [(stage input, its
    moment block), ...]) in stage order.  With need_out False the last stage
    (whose output the adjoint does not read) is skipped and f_next is None."""
    m0 = _moments_of(s, f)
    if s._stepper == 'RK2':
        uh = torch.empty_like(f)
        mh = torch.empty_like(m0)
        s._stage(f, f, None, uh, 0.0, 1.0, dt / 2, m0, mh)
        out = None
        if need_out:
            out = torch.empty_like(f)
            s._stage(uh, f, None, out, 0.0, 1.0, dt, mh, None)
        return out, [(f, m0), (uh, mh)]
    u1, m1 = torch.empty_like(f), torch.empty_like(m0)
    s._stage(f, f, None, u1, 0.0, 1.0, dt, m0, m1)
    u2, m2 = torch.empty_like(f), torch.empty_like(m0)
    s._stage(u1, u1, f, u2, 0.75, 0.25, dt, m1, m2)
    out = None
    if need_out:
        out = torch.empty_like(f)
        s._stage(u2, u2, f, out, 1 / 3, 2 / 3, dt, m2, None)
    return out, [(f, m0), (u1, m1), (u2, m2)]


def _adjoint_stage(s, u, mom, lam, base, a, b, c, ws, tau_q, base2=None, a2=0.0):
    out = torch.empty_like(lam)
    adj = _lib.Adjoint(_lib.ptr(u), _lib.ptr(mom), _lib.ptr(lam), _lib.ptr(base), _lib.ptr(out),
                       float(a), float(b), float(c), _lib.ptr(tau_q), _lib.ptr(ws),
                       _lib.ptr(base2), float(a2))
    _lib.check(s._L.boltx_adjoint_stage(ctypes.byref(s._grid), ctypes.byref(s._phys),
                                        ctypes.byref(adj), _lib.stream_handle()))
    return out


def _adjoint_step(s, tape, lam, dt, ws, tau_rows):
    """Reverse one step: lam (adjoint of the step's output) -> adjoint of its
    input; each stage's d/d tau per q-cell goes into the next row that
    ``tau_rows()`` hands out."""
    src = bool(s._source_on)

    def run(u, mom, lam_, base, a, b, c, base2=None, a2=0.0):
        return _adjoint_stage(s, u, mom, lam_, base, a, b, c, ws,
                              tau_rows() if src else None, base2, a2)

    if s._stepper == 'RK2':
        (x0, m0), (xh, mh) = tape
        # u^{n+1} = x0 + dt L(xh); xh = x0 + dt/2 L(x0)
        lh = run(xh, mh, lam, None, 0.0, 0.0, dt)
        return run(x0, m0, lh, lam, 1.0, 1.0, dt / 2)
    (x0, m0), (x1, m1), (x2, m2) = tape
    # u^{n+1} = 1/3 x0 + 2/3 (x2 + dt L(x2))
    l2 = run(x2, m2, lam, None, 0.0, 2 / 3, 2 / 3 * dt)
    # x2 = 3/4 x0 + 1/4 (x1 + dt L(x1))
    l1 = run(x1, m1, l2, None, 0.0, 0.25, 0.25 * dt)
    # x1 = x0 + dt L(x0); x0 also feeds u^{n+1} (1/3) and x2 (3/4)
    return run(x0, m0, l1, lam, 1 / 3, 1.0, dt, base2=l2, a2=0.75)


class _TauRows:
    """Rows of a [TAU_ROWS][Q] buffer for the stages' per-q-cell d/d tau:
    a full buffer is summed into the total (one reduction per TAU_ROWS
    stages, in a fixed order) and reused."""
    TAU_ROWS = 256

    def __init__(self, Q, device):
        self.buf = torch.empty((self.TAU_ROWS, Q), dtype=F64, device=device)
        self.total = torch.zeros((), dtype=F64, device=device)
        self.n = 0

    def __call__(self):
        if self.n == self.TAU_ROWS:
            self.flush()
        row = self.buf[self.n]
        self.n += 1
        return row

    def flush(self):
        if self.n:
            self.total.add_(self.buf[:self.n].sum())
        self.n = 0
        return self.total


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, tau, solver, n_steps, dt, every):
        s = solver
        _check_solver(s)
        _set_tau(s, float(tau))
        f = interior_to_internal(s, f0.detach())
        ckpts = []
        for k in range(n_steps):
            if k % every == 0:
                ckpts.append(f)
            f, _ = _step_recorded(s, f, dt)
        ctx.solver, ctx.ckpts = s, ckpts
        ctx.n_steps, ctx.dt, ctx.every, ctx.tau = n_steps, dt, every, float(tau)
        ctx.tau_shape, ctx.tau_dtype = tuple(tau.shape), tau.dtype
        s._f = f.clone()
        s._update_moments()
        return internal_to_interior(s, f)

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        _set_tau(s, ctx.tau)
        lam = interior_to_internal(s, g)
        nws = int(s._L.boltx_adjoint_workspace_size(ctypes.byref(s._grid),
                                                    ctypes.byref(s._phys)))
        ws = torch.empty(max(nws, 8), dtype=torch.uint8, device=s.device)
        tau_rows = _TauRows(s._Q, s.device)
        n, every, dt = ctx.n_steps, ctx.every, ctx.dt
        for seg in reversed(range(len(ctx.ckpts))):
            k0, k1 = seg * every, min(seg * every + every, n)
            f = ctx.ckpts[seg]
            tapes = []
            for k in range(k0, k1):
                # the segment's final state is the next checkpoint: not needed
                f, tape = _step_recorded(s, f, dt, need_out=k + 1 < k1)
                tapes.append(tape)
            for tape in reversed(tapes):
                lam = _adjoint_step(s, tape, lam, dt, ws, tau_rows)
        g_f0 = internal_to_interior(s, lam)
        g_tau = tau_rows.flush().to(ctx.tau_dtype).reshape(ctx.tau_shape)
        return g_f0, g_tau, None, None, None, None


def evolve(solver, f0, tau, n_steps, dt, checkpoint_every=10):
    """Advance the interior state ``f0`` ((P, N_q1, N_q2) float64 on the
    solver's device) by ``n_steps`` steps of ``dt`` with collision time
    ``tau`` (a float64 scalar tensor); differentiable in ``f0`` and ``tau``.
    The solver's own state is left at the final state."""
    if not torch.is_tensor(tau):
        tau = torch.tensor(float(tau), dtype=F64, device=f0.device)
    return _Evolve.apply(f0, tau, solver, int(n_steps), float(dt), int(checkpoint_every))


def moment(solver, f, name):
    """A built-in moment (moments.py:5-22: density, mom_v1_bulk, mom_v2_bulk,
    mom_v3_bulk, energy) of an interior state, shaped (N_q1, N_q2);
    differentiable (a weighted sum over p)."""
    v1 = solver.p1_center.reshape(-1)
    v2 = solver.p2_center.reshape(-1)
    v3 = solver.p3_center.reshape(-1)
    w = {'density': torch.ones_like(v1), 'mom_v1_bulk': v1, 'mom_v2_bulk': v2,
         'mom_v3_bulk': v3, 'energy': 0.5 * (v1 ** 2 + v2 ** 2 + v3 ** 2)}[name]
    dv = solver._dp[0] * solver._dp[1] * solver._dp[2]
    return torch.tensordot(w, f, dims=([0], [0])) * dv
