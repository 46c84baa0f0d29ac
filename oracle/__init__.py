"""CPU oracle of the LBcuda time step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product library never does.
"""
