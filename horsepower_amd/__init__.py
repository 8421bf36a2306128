"""MI355X (gfx950) HIP implementation of HorsePower's HorseIR primitive execution layer.

The compute path is libhorsepower_amd.so (HIP kernels behind the reference's pfn* C-ABI);
this package holds its build script, the ctypes binding, the seeded data generators and the
HorseIR query plans (TPC-H Q6/Q1/Q3 shapes) written against that ABI.
"""
from .vnode import *  # noqa: F401,F403
