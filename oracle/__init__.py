"""CPU oracle for the path-tracing hot path (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the product package.
"""
