"""Oracle: CPU restatements of SpeedyTransforms' transform! used as the checker.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product package.
"""
