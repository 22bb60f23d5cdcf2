"""Drop-in mirror of bolt/lib/physical_system.py:8-283.

Same constructor, same validation and the same derived attributes (dq, dp,
callbacks).  The reference also stores PETSc's MPI communicator and prints a
banner; this build runs one process per GPU and stores no communicator.
"""
import types


class physical_system(object):
    """physical_system.py:8-283: holds the domain, parameters and the user's
    initial conditions, advection terms, source and moments."""

    def __init__(self, domain, boundary_conditions, params, initial_conditions,
                 advection_term, source, moments):
        # physical_system.py:69-79
        for a in [a for a in dir(boundary_conditions) if not a.startswith('__')]:
            val = getattr(boundary_conditions, a)
            if not isinstance(val, (str, types.FunctionType, types.ModuleType, float)):
                raise TypeError('Expected attributes of boundary_conditions '
                                'to be of type string or functions')
        # :82-84
        if isinstance(initial_conditions, types.ModuleType) is False:
            raise TypeError('Expected initial_conditions to be of type function')
        # :87-89
        if isinstance(source, types.FunctionType) is False:
            raise TypeError('Expected source_or_sink to be of type function')
        # :92-99
        for a in [a for a in dir(advection_term) if not a.startswith('_')]:
            if not isinstance(getattr(advection_term, a),
                              (types.FunctionType, types.ModuleType)):
                raise TypeError('Expected attributes of advection_term '
                                'to be of type function or module')
        # :101-110
        for a in [a for a in dir(moments) if not a.startswith('_')]:
            if not isinstance(getattr(moments, a), (types.FunctionType, types.ModuleType,
                                                     tuple)):
                raise TypeError('Expected attributes of moment_defs '
                                'to be of type function or module')
        # :113-127
        N_species = len(params.charge)
        self.N_species = N_species
        if len(params.mass) != len(params.charge):
            raise Exception('Inconsistenty in number of species. Mismatch between '
                            'the number of species mentioned in charge and mass inputs')
        if getattr(params, 'hybrid_model_enabled', False) is True and N_species != 1:
            raise Exception('Only ions are to be initialized when using the hybrid model')
        # :129-144
        for name in ('p1_start', 'p2_start', 'p3_start', 'p1_end', 'p2_end', 'p3_end'):
            if type(getattr(domain, name)) != list:
                raise Exception('p-domain start and end points need to be specified in a list')
        if not (len(domain.p1_start) == len(domain.p2_start) == len(domain.p3_start)
                == len(domain.p1_end) == len(domain.p2_end) == len(domain.p3_end)
                == N_species):
            raise Exception('Inconsistency in definition of p-space resolution with '
                            'number of species')
        # :147-157
        self.N_q1, self.q1_start, self.q1_end = domain.N_q1, domain.q1_start, domain.q1_end
        self.N_q2, self.q2_start, self.q2_end = domain.N_q2, domain.q2_start, domain.q2_end
        self.N_p1, self.p1_start, self.p1_end = domain.N_p1, domain.p1_start, domain.p1_end
        self.N_p2, self.p2_start, self.p2_end = domain.N_p2, domain.p2_start, domain.p2_end
        self.N_p3, self.p3_start, self.p3_end = domain.N_p3, domain.p3_start, domain.p3_end
        # :160-174
        if (self.N_q1 < 0 or self.N_q2 < 0 or self.N_p1 < 0 or self.N_p2 < 0
                or self.N_p3 < 0 or domain.N_ghost < 0):
            raise Exception('Grid resolution for the phase space cannot be negative')
        if (self.q1_start > self.q1_end or self.q2_start > self.q2_end
                or self.p1_start > self.p1_end or self.p2_start > self.p2_end
                or self.p3_start > self.p3_end):
            raise Exception('Start point cannot be placed after the end point')
        # :177-187
        self.dq1 = (self.q1_end - self.q1_start) / self.N_q1
        self.dq2 = (self.q2_end - self.q2_start) / self.N_q2
        self.dp1, self.dp2, self.dp3 = [], [], []
        for i in range(N_species):
            self.dp1.append((self.p1_end[i] - self.p1_start[i]) / self.N_p1)
            self.dp2.append((self.p2_end[i] - self.p2_start[i]) / self.N_p2)
            self.dp3.append((self.p3_end[i] - self.p3_start[i]) / self.N_p3)
        # :191-213
        self.N_ghost = domain.N_ghost
        self.boundary_conditions = boundary_conditions
        self.params = params
        self.initial_conditions = initial_conditions
        self.A_q = advection_term.A_q
        self.C_q = advection_term.C_q
        self.A_p = advection_term.A_p
        self.C_p = advection_term.C_p
        self.source = source
        self.moments = moments
        # :216 (PETSc.COMM_WORLD) -- single process per GPU here
        self.mpi_communicator = None
        # :218-230
        if getattr(params, 'fields_type', None) == 'electrodynamic':
            if params.fields_solver.upper() != 'FDTD':
                raise Exception('Solver specified isn\'t an electrodynamic solver')
        if getattr(params, 'fields_type', None) == 'electrostatic':
            if params.fields_solver.upper() not in ('SNES', 'FFT'):
                raise Exception('Solver specified isn\'t an electrostatic solver')
