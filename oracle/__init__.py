"""Oracle: CPU restatement of the reference path tracer (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this package,
always as the checker. The fhp product never imports it.
"""
from .pyoracle import Oracle, OracleScene, build_oracle, oracle_lib, packed_precision  # noqa: F401
