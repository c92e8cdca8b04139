"""ORACLE — test infrastructure, never the product.

CPU restatement of the reference's algorithms (crest_oracle.c + oracle.py), used
only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
