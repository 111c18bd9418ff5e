"""CPU oracle of the volumetric light-transport path: TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
checker; the product path never imports it.
"""
