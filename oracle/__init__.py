"""CPU restatement of the reference's fused esoteric-pull step.

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
