"""``pysweep.core.solver.Solver`` (the reference's core/solver.py:9-189): the
MI355X solver of pysweep_amd under the reference's module path."""
from pysweep_amd.solver import Solver  # noqa: F401

__all__ = ["Solver"]
