"""``python -m pysweep_amd``: the command line (pysweep_amd.cli)."""
from pysweep_amd.cli import main

main()
