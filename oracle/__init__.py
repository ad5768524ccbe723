"""ORACLE — test infrastructure only: a CPU restatement of the reference's ES
step (refcpu.cpp) and its ctypes driver (refcpu.py). Imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product."""
