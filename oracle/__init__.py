"""FP64 CPU oracle for Bolt's nonlinear finite-volume time step.

TEST INFRASTRUCTURE ONLY.  Nothing in ``bolt_amd`` imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker (never as the measured or
shipped path).

The reference (QuazarTech/Bolt) is pure Python on ArrayFire + PETSc, neither of
which is installed here, so the oracle is a restatement: ``bolt_oracle`` mirrors
the reference's whole-array op sequence in torch FP64 on the CPU, with each
function citing the reference file:line it follows.  How each function is
pinned (closed forms, reference known-answer tests, invariants) is listed in the
module docstring of ``bolt_oracle``.
"""
