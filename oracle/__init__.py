"""ORACLE — test infrastructure only.

CPU restatements of NanoPyx's eSRRF path used to check the HIP implementation: an FP32 C/OpenMP
restatement of the reference's threaded chain (``oracle.ref32``) and an FP64 NumPy statement of the same
formulas (``oracle.fp64``). Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package; the product package ``nanopyx_amd`` never does.
"""
