"""CPU oracle (test infrastructure only): a NumPy/SciPy restatement of the reference PM path used by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product never imports it."""
