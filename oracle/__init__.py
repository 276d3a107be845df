"""CPU oracle for the Pauli-propagation branch step (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package.
"""
