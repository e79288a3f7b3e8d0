"""The reference's plugin pairs (pysweep/equations/__init__.py:1): for each
name a Python half (``<name>.py``: set_globals, step and the
initial-condition helpers) and a device half (``<name>.cu``) side by side in
this directory, so ``os.path.join(<this directory>, "euler.py")`` names a
plugin file as it does in the reference.  The files are links to the
shipped plugins of pysweep_amd.equations."""
from pysweep.equations import example, euler, heat, checker, half  # noqa: F401

__all__ = ["example", "euler", "heat", "checker", "half"]
