"""CPU oracle of the symmetric-BEM EEG forward path -- TEST INFRASTRUCTURE ONLY.

Restates the reference's algorithm (OpenMEEG 2.5.16) on the CPU so that the
HIP product path can be checked against it.  Only tests/, the smoke() check
and bench.py's cpu_baseline leg may import this package.
"""
