"""FP64 CPU oracle of the PySPH hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product (pysph_amd) never does.
"""
