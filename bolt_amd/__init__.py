"""bolt_amd: MI355X-native (gfx950 HIP) implementation of Bolt's nonlinear
finite-volume time step behind Bolt's physical_system / nonlinear_solver API.

    from bolt_amd.physical_system import physical_system
    from bolt_amd.nonlinear_solver import nonlinear_solver
    import bolt_amd.src.nonrelativistic_boltzmann.advection_terms as advection_terms
    import bolt_amd.src.nonrelativistic_boltzmann.collision_operator as collision_operator
    import bolt_amd.src.nonrelativistic_boltzmann.moments as moments
"""
__version__ = "0.1.0"
