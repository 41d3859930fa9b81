"""thindb_amd — an MI355X (gfx950) engine for thinDB's scan -> filter ->
GROUP BY -> top-N path, behind thinDB's grouped-operator interface.

Modules:
  gpu   — ctypes binding of libthindb_gpu.so (C-ABI: include/thindb_gpu.h)
  exec  — Request / GpuGroupQuery / Table: the operator interface
  gen   — seeded synthetic tables of the benchmark workloads
"""
from .gpu import AggOp, PredOp, Type  # noqa: F401
