"""The reference's ``pysweep.core`` namespace (core/solver.py)."""
