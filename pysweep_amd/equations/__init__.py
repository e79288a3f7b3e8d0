"""Equation plugin pairs shipped with pysweep_amd (<name>.py + <name>.cu)."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def path(name):
    """Absolute path of a shipped plugin file, e.g. path('euler.cu')."""
    return os.path.join(HERE, name)
