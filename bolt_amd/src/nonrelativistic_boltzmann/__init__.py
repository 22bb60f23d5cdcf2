"""Nonrelativistic Boltzmann plugins: mirrors of bolt/src/nonrelativistic_boltzmann.

Passing these objects to ``physical_system`` selects the fused HIP stage
kernel; they are also ordinary torch callables with the reference's signatures.
"""
