"""CPU oracle for the FDMAP time-stepping path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product.  See
oracle/fdm_oracle.h.
"""
