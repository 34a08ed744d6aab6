"""CPU FP64 oracle of MDFF's force/field evaluation -- test infrastructure
only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg)."""
