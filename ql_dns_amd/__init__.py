"""MI355X-native quasilinear shearing-box time step (HIP, gfx950).

The compute path is a C-ABI shared library (include/ql_dns.h) built from
ql_dns_amd/csrc; this package holds its ctypes binding, the run parameters
and the seeded initial-state generator.
"""
from .params import QLParams, workload, WORKLOADS  # noqa: F401
