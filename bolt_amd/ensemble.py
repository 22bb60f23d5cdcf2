"""Ensembles of independent one-dimensional-in-q problems advanced together.

The reference steps one problem per process (nonlinear_solver.py:67,
timestep.py:101-160); many small problems then cost one launch sequence each.
``Ensemble`` takes already built ``nonlinear_solver`` objects of uneven sizes
(periodic q1, N_q2 = 1, fields off, the same reconstruction and precision)
with a time step each, pools their states in the q1-line layout and advances
all of them with ``boltx_ensemble_step``: per RK stage one launch of a workgroup per
(member, chunk of whole q1 lines) that forms the stage and the FP64 partial
moments of its output, and one batched moment-block / Maxwellian-table
launch.

The work shape (cells per chunk, cells a workgroup stages in LDS at once) is
picked by ``autotune``, which times candidate shapes on a scratch copy of the
ensemble and caches the winner per ensemble signature, in the process and in
a JSON file (per GPU model) that later processes read instead of tuning
again.  The chunk size groups the moment sums, so it changes results in the
last bits; the batch size does not.
"""
import ctypes
import itertools
import json
import os
import time

import torch

from . import _lib

F64 = torch.float64

DEFAULT_TUNING = (16384, 2048)
TUNING_FIELDS = len(_lib.EnsTuning._fields_)
_TUNED = {}
# Tunings persist across processes in a JSON file (one entry per ensemble
# signature and GPU model), by default under the user's cache directory.
DEFAULT_TUNING_CACHE = os.path.join(os.environ.get('XDG_CACHE_HOME') or
                                    os.path.join(os.path.expanduser('~'), '.cache'),
                                    'bolt_amd', 'ensemble_tuning.json')


def cache_key(signature, device_name):
    """JSON key of a tuning: the GPU model and the ensemble signature."""
    return device_name + '|' + json.dumps(signature)


def load_tuning(path, key):
    """The cached tuning for ``key`` in the JSON file ``path``, or None (no
    file, an unreadable or malformed file, no entry, a malformed entry)."""
    try:
        with open(path) as fh:
            table = json.load(fh)
        t = table[key]
        if isinstance(t, list) and len(t) == TUNING_FIELDS and all(isinstance(x, int) for x in t):
            return tuple(t)
    except (OSError, ValueError, KeyError, TypeError):
        pass
    return None


def store_tuning(path, key, tuning):
    """Add ``key`` -> ``tuning`` to the JSON file ``path`` (written to a
    temporary file and renamed, so a reader never sees a partial file).
    Returns False if the file cannot be written."""
    try:
        try:
            with open(path) as fh:
                table = json.load(fh)
            if not isinstance(table, dict):
                table = {}
        except (OSError, ValueError):
            table = {}
        table[key] = [int(x) for x in tuning]
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = '%s.%d.tmp' % (path, os.getpid())
        with open(tmp, 'w') as fh:
            json.dump(table, fh, indent=1)
        os.replace(tmp, path)
        return True
    except OSError:
        return False


def _tuning(t):
    return _lib.EnsTuning(*[int(x) for x in t])


class Ensemble:
    def __init__(self, solvers, dts, tuning=None, autotune=True, tuning_cache=None):
        """``tuning_cache``: path of the JSON tuning cache (None: the default
        under the user's cache directory; False: none)."""
        if len(solvers) != len(dts) or not solvers:
            raise ValueError('one time step per solver')
        self.solvers = list(solvers)
        s0 = self.solvers[0]
        self._L = _lib.load()
        for s in self.solvers:
            if s.N_q2 != 1 or s._fields_on:
                raise NotImplementedError('ensemble members need N_q2 = 1 and fields off')
            if s._general:
                raise NotImplementedError('ensemble members use the built-in C_q and BGK '
                                          'with a uniform tau')
            if s._stepper != s0._stepper or s._precision != s0._precision \
                    or s._recon_q != s0._recon_q:
                raise NotImplementedError('ensemble members must share stepper, precision '
                                          'and reconstruction')
        self.device = s0.device
        self.dtype = s0._dtype
        self.stepper = _lib.STEPPER[s0._stepper]
        n = len(self.solvers)
        self._n = n
        self._members = (_lib.EnsMember * n)()
        for i, (s, dt) in enumerate(zip(self.solvers, dts)):
            self._members[i].grid = s._grid
            self._members[i].phys = s._phys
            self._members[i].dt = float(dt)
        offs = (ctypes.c_longlong * n)()
        total = self._L.boltx_ensemble_layout(n, self._members, offs)
        if total < 0:
            _lib.check(-1)
        self.offsets = [int(o) for o in offs]
        self.f = torch.empty(int(total), dtype=self.dtype, device=self.device)
        for i, s in enumerate(self.solvers):
            self._view(self.f, i).copy_(self._lines(s))
        self.u1 = torch.empty_like(self.f)
        self.u2 = torch.empty_like(self.f)
        self.time_elapsed = 0.0
        self.steps_taken = 0
        self.tuning_cache = DEFAULT_TUNING_CACHE if tuning_cache is None else tuning_cache
        key = self.signature()
        if tuning is None:
            tuning = _TUNED.get(key)
        if tuning is None and self.tuning_cache:
            tuning = load_tuning(self.tuning_cache, self._cache_key())
            if tuning is not None:
                _TUNED[key] = tuning
        if tuning is None:
            tuning = self.autotune() if autotune else DEFAULT_TUNING
        self._set_tuning(tuning)

    # -- layout --------------------------------------------------------------
    def _view(self, pooled, i):
        s = self.solvers[i]
        P = s._P
        return pooled[self.offsets[i]:self.offsets[i] + P * s.N_q1].view(P, s.N_q1)

    @staticmethod
    def _lines(s):
        """A solver's interior state as [P][N_q1] (q1-line order)."""
        f = s._f
        if s._layout == _lib.LAYOUT['q1-lines']:
            return f.reshape(s._P, s.N_q1)
        return f.reshape(s.N_q1, s._P).t()

    def _cache_key(self):
        return cache_key(self.signature(), torch.cuda.get_device_name(self.device))

    def signature(self):
        return tuple((s.N_q1, s.N_p1, s.N_p2, s.N_p3) for s in self.solvers) + \
            (self.stepper, int(self._members[0].phys.precision), int(self._members[0].phys.recon_q))

    def _set_tuning(self, tuning):
        self.tuning = tuple(int(x) for x in tuning)
        self._tune = _tuning(self.tuning)
        nws = int(self._L.boltx_ensemble_workspace_size(self._n, self._members,
                                                        ctypes.byref(self._tune)))
        if nws == 0:
            _lib.check(-1)
        self._ws = torch.empty(nws, dtype=torch.uint8, device=self.device)
        mo, fo = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(self._L.boltx_ensemble_map(self._n, self._members, ctypes.byref(self._tune),
                                              ctypes.byref(mo), ctypes.byref(fo)))
        self._mom_off, self._flag_off = mo.value, fo.value
        _lib.check(self._L.boltx_ensemble_init(self._n, self._members, ctypes.byref(self._tune),
                                               _lib.ptr(self.f), _lib.ptr(self._ws),
                                               _lib.stream_handle()))

    # -- stepping ------------------------------------------------------------
    def step(self, n_steps=1):
        """Advance every member by n_steps of its own dt (no host sync)."""
        _lib.check(self._L.boltx_ensemble_step(
            self._n, self._members, ctypes.byref(self._tune), self.stepper, _lib.ptr(self.f),
            _lib.ptr(self.u1), _lib.ptr(self.u2), _lib.ptr(self._ws), int(n_steps),
            _lib.stream_handle()))
        self.steps_taken += n_steps

    def capture(self, n_steps):
        """Record n_steps steps into a HIP graph (torch.cuda.CUDAGraph); each
        ``replay()`` then advances every member by n_steps with one launch of
        the graph.  The captured kernels use no memsets and the work lists do
        not change, so replays repeat the eager launches exactly."""
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            _lib.check(self._L.boltx_ensemble_step(
                self._n, self._members, ctypes.byref(self._tune), self.stepper,
                _lib.ptr(self.f), _lib.ptr(self.u1), _lib.ptr(self.u2), _lib.ptr(self._ws),
                int(n_steps), _lib.stream_handle()))
        self._graph, self._graph_steps = g, int(n_steps)
        return g

    def replay(self):
        """Run the captured steps (see capture)."""
        self._graph.replay()
        self.steps_taken += self._graph_steps

    def check_divergence(self):
        """timestep.py:28-38 for every member (one host sync)."""
        flag = self._ws[self._flag_off:self._flag_off + 4].view(torch.int32)
        if int(flag.item()) != 0:
            raise SystemExit('Solver Diverging!')

    # -- results -------------------------------------------------------------
    def state(self, i):
        """Member i's interior f as (P, N_q1) in the storage dtype (a view)."""
        return self._view(self.f, i)

    def moments(self, i):
        """Member i's moment block [12][N_q1] (FP64 view): density, the three
        bulk momenta and energy (times dp1 dp2 dp3), then the BGK parameters."""
        off = self._mom_off // 8 + sum(_lib.MOM_ROWS * s.N_q1 for s in self.solvers[:i])
        n1 = self.solvers[i].N_q1
        flat = self._ws[:(self._ws.numel() // 8) * 8].view(F64)
        return flat[off:off + _lib.MOM_ROWS * n1].view(_lib.MOM_ROWS, n1)

    def write_back(self):
        """Copy every member's state into its solver (solver.f then reads it)."""
        for i, s in enumerate(self.solvers):
            lines = self.state(i)
            if s._layout == _lib.LAYOUT['q1-lines']:
                s._f.copy_(lines.reshape(s._f.shape))
            else:
                s._f.copy_(lines.t().reshape(s._f.shape))
            s._update_moments()

    # -- autotuner -----------------------------------------------------------
    def autotune(self, candidates=None, steps=2):
        """Time each candidate work shape (chunk_cells, batch_cells) over
        ``steps`` steps on a scratch copy of the ensemble and keep the
        fastest; the result is cached for ensembles with the same signature.
        Returns the chosen tuning."""
        if candidates is None:
            candidates = list(itertools.product((4096, 8192, 16384, 32768, 65536),
                                                (1024, 2048)))
        saved_f = self.f
        best, best_t = None, float('inf')
        for cand in candidates:
            self.f = saved_f.clone()
            try:
                self._set_tuning(cand)
            except _lib.BoltxError:
                continue
            self.step(1)  # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.step(steps)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if t < best_t:
                best, best_t = tuple(cand), t
        self.f = saved_f
        self.steps_taken = 0
        _TUNED[self.signature()] = best
        if self.tuning_cache:
            store_tuning(self.tuning_cache, self._cache_key(), best)
        self.tune_time_per_step = best_t / steps
        return best
