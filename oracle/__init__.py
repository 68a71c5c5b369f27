"""Correctness oracle (TEST INFRASTRUCTURE ONLY): FP64 NumPy restatement of the reference
semantics (oracle.py) and an FP32 op-order-pinned C twin (twin.c / twin.py).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it; the product never does."""
