"""Oracle: CPU restatements of SpeedyTransforms' transform!, the parity reference.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product package.
"""
