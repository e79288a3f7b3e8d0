"""``python -m pysweep``: the reference's ``pysweep`` console command."""
from pysweep.utils.commandline import commandLine

commandLine()
