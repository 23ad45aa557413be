"""MI355X-native Gaussian-mixture volumetric path tracer.

Python side of the package: a ctypes binding of the HIP library
(volprim.py), and the synthetic scene generators of the benchmark
configurations (scenes.py).  The directory name is not a Python identifier;
load it with volprim_amd_loader() below or importlib, as tests/ and bench.py do.
"""
