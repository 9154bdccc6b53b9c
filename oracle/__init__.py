"""Test infrastructure only: the CPU oracle (a C restatement of the reference renderer) and its
ctypes binding. Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg."""
