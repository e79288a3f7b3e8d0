"""``pysweep`` command line (the reference's utils/commandline.py:28-144):
``pysweep -f euler|heat -nx N -nt T -b B -s SHARE [--swept]``, run by
``commandLine()`` (the console entry point) or ``python -m pysweep``.  The
options and problem set-up are pysweep_amd.cli's."""
import os

from pysweep_amd.cli import adjust_array_size, command_line, run_euler, run_heat, str2bool  # noqa: F401

eqnPath = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "equations")


def getEqnPath(eqn):
    """Path of a shipped plugin file, e.g. getEqnPath("euler.cu")
    (commandline.py:8-10)."""
    return os.path.join(eqnPath, eqn)


# the reference's names for the same functions
adjustArraySize = adjust_array_size
runEuler = run_euler
runHeat = run_heat


def commandLine(argv=None):
    """Parse the command line and run the chosen example (commandline.py:117-144)."""
    import sys
    return command_line(sys.argv[1:] if argv is None else argv)
