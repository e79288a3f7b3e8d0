"""pysweep_amd: the swept space-time decomposition of PySweep rebuilt for one
AMD MI355X (gfx950).

``Solver`` keeps the reference's interface (pysweep.Solver); plugins are the
reference's pairs of a Python module (set_globals / step) and a device source
(step(double*, int, int) with __constant__ globals).  The phase kernels run
from LDS-resident tiles compiled with hiprtc for gfx950 and are driven
through the C ABI of libpysweep_hip.so (include/pysweep_hip.h).
"""
from pysweep_amd.solver import Solver  # noqa: F401
from pysweep_amd import equations  # noqa: F401

__all__ = ["Solver", "equations"]
