"""Test infrastructure: the CPU oracle for the 2D CSR wake step.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product (pydfcsr_amd) never does.
"""
