"""``pysweep``: the reference's import surface on top of pysweep_amd.

Reference callers keep their code unchanged::

    import pysweep
    solver = pysweep.Solver("eulerConditions.hdf5", "euler.yaml")
    solver()

Exports as the reference's package (pysweep/__init__.py:1-5): ``Solver``,
the plugin pairs ``equations.{example, euler, heat, checker, half}`` (each a
Python half with set_globals / step and its initial-condition helpers, and a
device half ``<name>.cu`` in the same directory) and ``utils.commandline``
(the ``pysweep`` command line).  The solver runs on one MI355X through the
HIP kernels of pysweep_amd; nothing here computes on the host except the
plugins' own CPU halves.
"""
from pysweep.core.solver import Solver  # noqa: F401
from pysweep import equations  # noqa: F401
from pysweep import utils  # noqa: F401

__all__ = ["Solver", "equations", "utils"]
