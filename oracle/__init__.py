"""ORACLE -- test infrastructure only: the C restatement of the reference CPU
emulator and its ctypes loader.  Imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg only."""
