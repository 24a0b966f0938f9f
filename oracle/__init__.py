"""CPU oracle (test infrastructure): a serial C restatement of the reference hot path and its
Python bindings. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it."""
