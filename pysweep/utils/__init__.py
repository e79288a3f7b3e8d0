"""The reference's ``pysweep.utils`` namespace (utils/commandline.py)."""
from pysweep.utils import commandline  # noqa: F401

__all__ = ["commandline"]
