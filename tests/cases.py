"""Problem definitions shared by the tests, the smoke check and the bench.

Each case is built as the reference's user modules (domain, boundary
conditions, params, initialize) with torch callables, so the same modules drive
the oracle (CPU tensors) and bolt_amd (GPU tensors).  Workloads follow the
synthetic Maxwellian-perturbation problems: 1D1V minmod RK2, 2D2V WENO5 SSP-RK3,
2D3V electromagnetic, at sizes the oracle finishes in seconds unless stated.
"""
import math
import types

import numpy as np
import torch


def _module(name, **attrs):
    m = types.ModuleType(name)
    for k, v in attrs.items():
        setattr(m, k, v)
    return m


def _const_tau(value):
    def tau(q1, q2, p1, p2, p3):
        return value * (q1 * 0 + 1) * (p1 * 0 + 1)
    return tau


def maxwellian_case(*, N_q1, N_q2, N_p1, N_p2=1, N_p3=1, N_ghost=3, p_max=10.0,
                    p_dim=1, modes=((1, 0),), amps=(1e-3,), phases=(0.0,),
                    recon_q="weno5", recon_p="weno5", riemann="upwind-flux",
                    stepper="SSP-RK3", tau=0.01, source=True, fields=False,
                    B0=0.5, charge=-1.0, mass=1.0, precision="fp64",
                    q2_bc="periodic", p3_max=None, memory_layout="auto",
                    fields_type="electrodynamic", fields_initialize="user-defined"):
    """Maxwellian (T = 1, u = 0) with density n = 1 + sum A_k cos(2 pi (kx q1 +
    ky q2) + phi_k) on [0,1)^2, p in [-p_max, p_max] per active axis (singleton
    axes span [-0.5, 0.5] as in the reference examples)."""
    def ext(N, pm):
        return ([-0.5], [0.5]) if N == 1 else ([-pm], [pm])
    p1s, p1e = ext(N_p1, p_max)
    p2s, p2e = ext(N_p2, p_max)
    p3s, p3e = ext(N_p3, p3_max or p_max)
    domain = _module("domain", q1_start=0.0, q1_end=1.0, N_q1=N_q1,
                     q2_start=0.0, q2_end=1.0, N_q2=N_q2,
                     p1_start=p1s, p1_end=p1e, N_p1=N_p1,
                     p2_start=p2s, p2_end=p2e, N_p2=N_p2,
                     p3_start=p3s, p3_end=p3e, N_p3=N_p3, N_ghost=N_ghost)
    bc = _module("boundary_conditions", in_q1_left="periodic", in_q1_right="periodic",
                 in_q2_bottom=q2_bc, in_q2_top=q2_bc)
    params = _module(
        "params", fields_type=fields_type, fields_initialize=fields_initialize,
        fields_solver="fdtd" if fields_type == "electrodynamic" else "fft",
        solver_method_in_q="FVM", solver_method_in_p="FVM",
        reconstruction_method_in_q=recon_q, reconstruction_method_in_p=recon_p,
        riemann_solver_in_q=riemann, riemann_solver_in_p=riemann, p_dim=p_dim,
        num_devices=1, mass=[mass], boltzmann_constant=1, charge=[charge],
        fields_enabled=fields, source_enabled=source, instantaneous_collisions=False,
        hybrid_model_enabled=False, eps=1, mu=1, tau=_const_tau(tau),
        time_stepper_fvm=stepper, precision=precision, B0=B0, memory_layout=memory_layout,
        modes=tuple(modes), amps=tuple(amps), phases=tuple(phases))

    def initialize_f(q1, q2, p1, p2, p3, params):
        n = 1.0 + 0.0 * q1
        for (kx, ky), A, ph in zip(params.modes, params.amps, params.phases):
            n = n + A * torch.cos(2 * math.pi * (kx * q1 + ky * q2) + ph)
        m, k, T = params.mass, params.boltzmann_constant, 1.0
        f = n * (m / (2 * math.pi * k * T)) ** (params.p_dim / 2) \
            * torch.exp(-m * p1**2 / (2 * k * T))
        if params.p_dim >= 2:
            f = f * torch.exp(-m * p2**2 / (2 * k * T))
        if params.p_dim >= 3:
            f = f * torch.exp(-m * p3**2 / (2 * k * T))
        return f

    def initialize_E(q1, q2, params):
        z = 0.0 * q1
        return (z, z, z)

    def initialize_B(q1, q2, params):
        z = 0.0 * q1
        return (z, z, z + params.B0)

    init = _module("initialize", initialize_f=initialize_f, initialize_E=initialize_E,
                   initialize_B=initialize_B)
    return types.SimpleNamespace(domain=domain, boundary_conditions=bc, params=params,
                                 initialize=init)


def draw(seed, lo, hi, n=None):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, n)


def case_c1(seed=0, **over):
    """1D1V, fields off, minmod, upwind, BGK tau = 0.01, RK2, dt = 1e-3:
    N_q1 = 32 (2 ghosts), N_p1 = 32 on [-10, 10]."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(1e-4, 1e-3)
    phi = rng.uniform(0, 2 * math.pi)
    kw = dict(N_q1=32, N_q2=1, N_p1=32, N_ghost=2, p_max=10.0, p_dim=1,
              modes=((1, 0),), amps=(A,), phases=(phi,), recon_q="minmod",
              recon_p="minmod", riemann="upwind-flux", stepper="RK2", tau=0.01)
    kw.update(over)
    c = maxwellian_case(**kw)
    c.dt, c.steps = 1e-3, 100
    return c


def case_2d2v(seed=0, N_q=128, N_p=48, **over):
    """2D2V, fields off, WENO5, LLF, BGK tau = 1e-2, SSP-RK3, dt = 5e-4: the
    full size is N_q = 128, N_p = 48 on [-10, 10]^2 (3 ghosts)."""
    rng = np.random.default_rng(seed)
    modes = ((1, 1), (1, 2), (2, 1), (2, 2))
    A = rng.uniform(1e-4, 1e-3, 4)
    phi = rng.uniform(0, 2 * math.pi, 4)
    kw = dict(N_q1=N_q, N_q2=N_q, N_p1=N_p, N_p2=N_p, N_ghost=3, p_max=10.0, p_dim=2,
              modes=modes, amps=tuple(A), phases=tuple(phi), recon_q="weno5",
              recon_p="weno5", riemann="lax-friedrichs", stepper="SSP-RK3", tau=1e-2)
    kw.update(over)
    c = maxwellian_case(**kw)
    c.dt, c.steps = 5e-4, 200
    return c


def case_2d3v(seed=0, N_q=64, N_p=24, **over):
    """2D3V electromagnetic: electrons (q/m = -1), PPM in q, WENO5 in p, LLF,
    BGK tau = 0.1, p in [-8, 8]^3, B = (0, 0, 0.5), E = 0, dt = 2e-4."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(1e-3, 1e-2)
    phi = rng.uniform(0, 2 * math.pi)
    kw = dict(N_q1=N_q, N_q2=N_q, N_p1=N_p, N_p2=N_p, N_p3=N_p, N_ghost=3, p_max=8.0,
              p_dim=3, modes=((1, 1),), amps=(A,), phases=(phi,), recon_q="ppm",
              recon_p="weno5", riemann="lax-friedrichs", stepper="SSP-RK3", tau=0.1,
              fields=True, B0=0.5, charge=-1.0)
    kw.update(over)
    c = maxwellian_case(**kw)
    c.dt, c.steps = 2e-4, 500
    return c


def oracle_solver(case):
    """The oracle for a case; a case's own ``advection`` module and
    ``source`` function (user plugins written with torch ops) replace the
    built-in ones."""
    from oracle import bolt_oracle as bo
    sysm = bo.make_system(case.domain, case.boundary_conditions, case.params, case.initialize,
                          getattr(case, "advection", None) or bo.advection_terms,
                          getattr(case, "source", None) or bo.BGK, bo.moments)
    return bo.OracleSolver(sysm)


def hip_solver(case):
    from bolt_amd.physical_system import physical_system
    from bolt_amd.nonlinear_solver import nonlinear_solver
    import bolt_amd.src.nonrelativistic_boltzmann.advection_terms as adv
    import bolt_amd.src.nonrelativistic_boltzmann.collision_operator as coll
    import bolt_amd.src.nonrelativistic_boltzmann.moments as mom
    system = physical_system(case.domain, case.boundary_conditions, case.params,
                             case.initialize, getattr(case, "advection", None) or adv,
                             getattr(case, "source", None) or coll.BGK, mom)
    return nonlinear_solver(system)


def smoothed_noise(seed, N1, N2, sigma=2.0, amp=1e-3):
    """Gaussian-smoothed N(0, 1) noise on the periodic N1 x N2 grid (FFT
    convolution, sigma in cells), rescaled to max |dn| = amp and stored in
    FP32, returned as float64."""
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((N1, N2)))
    k1 = torch.fft.fftfreq(N1).reshape(-1, 1)
    k2 = torch.fft.fftfreq(N2).reshape(1, -1)
    g = torch.exp(-2 * (math.pi * sigma) ** 2 * (k1 ** 2 + k2 ** 2))
    y = torch.fft.ifft2(torch.fft.fft2(x) * g).real
    y = y * (amp / y.abs().max())
    return y.float().double()


def case_c4(seed=0, N_q=64, N_p=32, **over):
    """Adjoint workload: 2D2V fields off, WENO5 + LLF + BGK tau = 0.05,
    SSP-RK3, dt = 5e-4, 100 steps, p in [-10, 10]^2; the initial state is
    (1 + dn(q)) times the unit Maxwellian with dn = smoothed_noise."""
    kw = dict(N_q1=N_q, N_q2=N_q, N_p1=N_p, N_p2=N_p, N_ghost=3, p_max=10.0, p_dim=2,
              modes=((1, 1),), amps=(0.0,), phases=(0.0,), recon_q="weno5",
              recon_p="weno5", riemann="lax-friedrichs", stepper="SSP-RK3", tau=0.05)
    kw.update(over)
    c = maxwellian_case(**kw)
    c.dt, c.steps, c.checkpoint_every = 5e-4, 100, 10
    c.dn = smoothed_noise(seed, kw["N_q1"], kw["N_q2"])
    return c


def maxwellian_interior(p1, p2, p3, p_dim, dn):
    """(1 + dn(q)) * (2 pi)^(-p_dim/2) exp(-|p|^2/2) as a (P, N_q1, N_q2)
    tensor; p1..p3 are the (P,) cell-centre velocities."""
    e = p1 ** 2
    if p_dim >= 2:
        e = e + p2 ** 2
    if p_dim >= 3:
        e = e + p3 ** 2
    M = (2 * math.pi) ** (-p_dim / 2) * torch.exp(-0.5 * e)
    return M.reshape(-1, 1, 1) * (1 + dn).unsqueeze(0)


def case_c5(seed=0, n=96, precision="fp32-bulk", **over):
    """Ensemble workload: n independent 1D2V members (periodic q1 in [0, 1),
    N_q2 = 1, p in [-10, 10]^2), each drawn from the seed: N_q1 in
    {32, 64, 128, 256}, N_p1 = N_p2 in {24, 32, 48, 64}, tau log-uniform in
    [1e-3, 1], density mode k in {1..4} with A ~ U[1e-4, 1e-2]; Maxwellian
    T = 1, WENO5 + LLF + BGK, SSP-RK3 at dt = 0.4 dq1 / max|v1|.  Returns a
    list of (case, dt)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nq = int(rng.choice([32, 64, 128, 256]))
        npp = int(rng.choice([24, 32, 48, 64]))
        tau = float(10 ** rng.uniform(-3, 0))
        k = int(rng.integers(1, 5))
        A = float(rng.uniform(1e-4, 1e-2))
        phi = float(rng.uniform(0, 2 * math.pi))
        kw = dict(N_q1=nq, N_q2=1, N_p1=npp, N_p2=npp, N_ghost=3, p_max=10.0, p_dim=2,
                  modes=((k, 0),), amps=(A,), phases=(phi,), recon_q="weno5", recon_p="weno5",
                  riemann="lax-friedrichs", stepper="SSP-RK3", tau=tau, precision=precision)
        kw.update(over)
        c = maxwellian_case(**kw)
        vmax = 10.0 - 0.5 * 20.0 / kw["N_p1"]
        dt = 0.4 * (1.0 / kw["N_q1"]) / vmax
        c.dt, c.steps = dt, 200
        out.append((c, dt))
    return out


def case_es(seed=0, N_q1=32, N_q2=1, N_p=24, **over):
    """Electrostatic (FFT Poisson) problem: electrons (q/m = -1) on a
    neutralising background, density mode (1, 0) (and (1, 1) in 2D) with
    A ~ U[1e-3, 1e-2], p in [-6, 6] per active axis, minmod in q and p, LLF,
    BGK tau = 0.1, RK2 (the reference stepper), dt = 1e-3."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(1e-3, 1e-2)
    phi = rng.uniform(0, 2 * math.pi)
    two_d = N_q2 > 1
    kw = dict(N_q1=N_q1, N_q2=N_q2, N_p1=N_p, N_p2=N_p if two_d else 1, N_ghost=3,
              p_max=6.0, p_dim=2 if two_d else 1,
              modes=((1, 1),) if two_d else ((1, 0),), amps=(A,), phases=(phi,),
              recon_q="minmod", recon_p="minmod", riemann="lax-friedrichs", stepper="RK2",
              tau=0.1, fields=True, charge=-1.0, B0=0.0, fields_type="electrostatic",
              fields_initialize="fft")
    kw.update(over)
    c = maxwellian_case(**kw)
    c.dt, c.steps = 1e-3, 100
    return c


def drifting_maxwellian(p1, p2, p3, p_dim, n, u1, u2, T=1.0):
    """n (2 pi T)^(-p_dim/2) exp(-((p1-u1)^2 + (p2-u2)^2 [+ p3^2]) / 2T)."""
    e = (p1 - u1) ** 2
    if p_dim >= 2:
        e = e + (p2 - u2) ** 2
    if p_dim >= 3:
        e = e + p3 ** 2
    return n * (2 * math.pi * T) ** (-p_dim / 2) * torch.exp(-e / (2 * T))


def case_bc(*, left="mirror", right="mirror", bottom="periodic", top="periodic",
            N_q1=8, N_q2=6, N_p1=6, N_p2=6, p_max=6.0, p_dim=2, N_ghost=3,
            recon_q="weno5", riemann="upwind-flux", stepper="RK2", source=False,
            tau=0.05, precision="fp64", drift=(0.7, -0.4), bump=0.3, inflow=None,
            dt=2e-3, electrostatic=False):
    """Non-periodic q boundaries (boundaries.py): a drifting Maxwellian with a
    Gaussian density bump centred in [0,1)^2 (width 0.2, amplitude ``bump``),
    mean velocity ``drift``, and per-side q-BCs; dirichlet sides inject a
    drifting Maxwellian (density, u1, u2) = ``inflow`` or (1.2, 0.5, 0.3)
    through f_left / f_right / f_bottom / f_top(f, t, q1, q2, p1, p2, p3,
    params).  1-D problems (N_q2 = 1) take bottom = top = 'none'.  With
    ``electrostatic`` the FFT Poisson fields of the electrons (charge -1)
    drive p-advection."""
    if N_q2 == 1:
        bottom = top = "none"
    es = dict(fields=True, fields_type="electrostatic", fields_initialize="fft") \
        if electrostatic else {}
    c = maxwellian_case(N_q1=N_q1, N_q2=N_q2, N_p1=N_p1, N_p2=N_p2 if p_dim >= 2 else 1,
                        N_ghost=N_ghost, p_max=p_max, p_dim=p_dim, recon_q=recon_q,
                        recon_p=recon_q, riemann=riemann, stepper=stepper, tau=tau,
                        source=source, precision=precision, memory_layout="auto", **es)
    n_in, u1_in, u2_in = inflow or (1.2, 0.5, 0.3)

    def f_side(f, t, q1, q2, p1, p2, p3, params):
        return drifting_maxwellian(p1, p2, p3, params.p_dim, n_in, u1_in, u2_in) \
            * (q1 * 0 + 1) * (q2 * 0 + 1)

    c.boundary_conditions = _module(
        "boundary_conditions", in_q1_left=left, in_q1_right=right,
        in_q2_bottom=bottom, in_q2_top=top, f_left=f_side, f_right=f_side,
        f_bottom=f_side, f_top=f_side)
    c.params.drift, c.params.bump = tuple(drift), bump

    def initialize_f(q1, q2, p1, p2, p3, params):
        r2 = (q1 - 0.5) ** 2 + ((q2 - 0.5) ** 2 if N_q2 > 1 else 0 * q2)
        n = 1.0 + params.bump * torch.exp(-r2 / (2 * 0.2 ** 2))
        return drifting_maxwellian(p1, p2, p3, params.p_dim, n, params.drift[0],
                                   params.drift[1])

    c.initialize.initialize_f = initialize_f
    c.dt, c.steps = dt, 4
    return c


def case_species(*, S=2, masses=(1.0, 4.0), charges=(-1.0, 1.0), temps=(1.0, 0.5),
                 drifts=(0.3, -0.2), p_max=(6.0, 3.0), taus=(0.05, 0.2), N_q1=16, N_q2=6,
                 N_p1=8, N_p2=6, N_p3=1, p_dim=2, fields=False, fields_type="electrostatic",
                 recon_q="weno5", recon_p="minmod", riemann="upwind-flux", stepper="RK2",
                 source=True, precision="fp64", memory_layout="auto", amp=3e-2, B0=0.2,
                 dt=1e-3):
    """S species on their own p-grids ([-p_max_s, p_max_s] per active axis),
    each a drifting Maxwellian with its own mass, temperature and collision
    time, density 1 + amp cos(2 pi (q1 + q2)) (shifted in phase per species),
    q periodic.  Fields optional: electrostatic (FFT) or electrodynamic
    (FDTD, B = (0, 0, B0))."""
    p_dim = min(p_dim, 3)
    pm = list(p_max[:S])

    def ext(N, k):
        return ([-0.5] * S, [0.5] * S) if N == 1 else ([-v for v in pm], list(pm))
    p1s, p1e = ext(N_p1, 0)
    p2s, p2e = ext(N_p2, 1)
    p3s, p3e = ext(N_p3, 2)
    domain = _module("domain", q1_start=0.0, q1_end=1.0, N_q1=N_q1, q2_start=0.0,
                     q2_end=1.0, N_q2=N_q2, p1_start=p1s, p1_end=p1e, N_p1=N_p1,
                     p2_start=p2s, p2_end=p2e, N_p2=N_p2, p3_start=p3s, p3_end=p3e,
                     N_p3=N_p3, N_ghost=3)
    q2_bc = "periodic" if N_q2 > 1 else "none"
    bc = _module("boundary_conditions", in_q1_left="periodic", in_q1_right="periodic",
                 in_q2_bottom=q2_bc, in_q2_top=q2_bc)
    tau_v = list(taus[:S])

    def tau(q1, q2, p1, p2, p3):
        t = torch.tensor(tau_v, dtype=torch.float64, device=p1.device).reshape(1, S, 1, 1)
        return t * (q1 * 0 + 1) * (p1 * 0 + 1)

    params = _module(
        "params", fields_type=fields_type,
        fields_initialize="fft" if fields_type == "electrostatic" else "user-defined",
        fields_solver="fft" if fields_type == "electrostatic" else "fdtd",
        solver_method_in_q="FVM", solver_method_in_p="FVM",
        reconstruction_method_in_q=recon_q, reconstruction_method_in_p=recon_p,
        riemann_solver_in_q=riemann, riemann_solver_in_p=riemann, p_dim=p_dim,
        num_devices=1, mass=list(masses[:S]), boltzmann_constant=1, charge=list(charges[:S]),
        fields_enabled=fields, source_enabled=source, instantaneous_collisions=False,
        hybrid_model_enabled=False, eps=1, mu=1, tau=tau, time_stepper_fvm=stepper,
        precision=precision, memory_layout=memory_layout, B0=B0, temps=list(temps[:S]),
        drifts=list(drifts[:S]), amp=amp)

    def initialize_f(q1, q2, p1, p2, p3, params):
        m = params.mass
        ph = torch.arange(S, dtype=torch.float64, device=p1.device).reshape(1, S, 1, 1)
        T = torch.tensor(params.temps, dtype=torch.float64, device=p1.device).reshape(1, S, 1, 1)
        u = torch.tensor(params.drifts, dtype=torch.float64, device=p1.device).reshape(1, S, 1, 1)
        n = 1.0 + params.amp * torch.cos(2 * math.pi * (q1 + q2) + 0.7 * ph)
        e = (p1 - u) ** 2
        if params.p_dim >= 2:
            e = e + p2 ** 2
        if params.p_dim >= 3:
            e = e + p3 ** 2
        return n * (m / (2 * math.pi * T)) ** (params.p_dim / 2) * torch.exp(-m * e / (2 * T))

    def initialize_E(q1, q2, params):
        z = 0.0 * q1
        return (z, z, z)

    def initialize_B(q1, q2, params):
        z = 0.0 * q1
        return (z, z, z + params.B0)

    init = _module("initialize", initialize_f=initialize_f, initialize_E=initialize_E,
                   initialize_B=initialize_B)
    c = types.SimpleNamespace(domain=domain, boundary_conditions=bc, params=params,
                              initialize=init)
    c.dt, c.steps = dt, 10
    return c


def single_species_of(case, k):
    """Species k of a multi-species case as a one-species problem."""
    def clone(m):
        return _module(m.__name__, **{a: v for a, v in vars(m).items()
                                      if not a.startswith("__")})
    c = types.SimpleNamespace(domain=clone(case.domain), params=clone(case.params),
                              boundary_conditions=case.boundary_conditions,
                              initialize=clone(case.initialize), dt=case.dt, steps=case.steps)
    d, p = c.domain, c.params
    for a in ("p1_start", "p1_end", "p2_start", "p2_end", "p3_start", "p3_end"):
        setattr(d, a, [getattr(d, a)[k]])
    for a in ("mass", "charge", "temps", "drifts"):
        setattr(p, a, [getattr(p, a)[k]])
    tv = case.params.tau(torch.zeros(1, 1, 1, 1, dtype=torch.float64),
                         torch.zeros(1, 1, 1, 1, dtype=torch.float64),
                         torch.zeros(1, len(case.params.mass), 1, 1, dtype=torch.float64),
                         None, None)[0, k].reshape(-1)[0].item()
    p.tau = lambda q1, q2, p1, p2, p3, _t=tv: _t * (q1 * 0 + 1) * (p1 * 0 + 1)
    f_all = case.initialize.initialize_f

    def initialize_f(q1, q2, p1, p2, p3, params, _k=k):
        # the multi-species initial condition with this species in column 0
        S = len(case.params.mass)
        pad = [torch.cat([x] * S, dim=1) if x is not None and x.shape[1] == 1 else x
               for x in (p1, p2, p3)]
        full_params = types.SimpleNamespace(**vars(case.params))
        full_params.mass = torch.tensor(case.params.mass, dtype=torch.float64,
                                        device=p1.device).reshape(1, S, 1, 1)
        return f_all(q1, q2, *pad, full_params)[:, _k:_k + 1]

    c.initialize.initialize_f = initialize_f
    return c
