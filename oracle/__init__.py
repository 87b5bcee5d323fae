"""CPU oracle for the pychem integral + Fock-build path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  It is the checker, never the thing measured or shipped.
"""
