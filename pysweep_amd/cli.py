"""Command line: ``python -m pysweep_amd -f euler|heat -nx N -nt T -b B [--swept]``.

Mirrors the reference's ``pysweep`` entry point (utils/commandline.py:28-144)
on top of ``Solver``: the same options and problem set-up (Euler: vortex on
[-5, 5]^2 sampled at ``linspace(-5, 5, nx)``, dx = 10/(nx-1), dt = 0.1 dx,
gamma = 1.4, k = 2, s = 2; heat: the decaying mode on ``linspace(0, 1, nx)``,
dx = 1/(nx-1), dt = 0.1 dx^2 / alpha, alpha = 1, k = 1, s = 1).  The
reference writes the initial condition to an HDF5 file with h5py's default
float32 dataset and reads it back; here it is built in memory and rounded to
float32 the same way.  Each run appends a record to ``log.yaml`` (as
core/io.py:10-22 does), with the device solve time and GCUPS added.

Differences: ``--share`` defaults to 1 (the whole grid on the GPU; below 1
the host takes the remaining block rows, swept only), the figure options
(-p) are not provided, and ``--exact`` takes exactly nt steps instead of
the reference's quantization to whole swept cycles.
"""
import argparse
import sys
import warnings

import numpy

from pysweep_amd import equations


_TRUTH = {"yes": True, "true": True, "t": True, "y": True, "1": True,
          "no": False, "false": False, "f": False, "n": False, "0": False}


def str2bool(v):
    """Boolean option values as the reference's CLI accepts them
    (utils/commandline.py:12-20)."""
    if isinstance(v, bool):
        return v
    try:
        return _TRUTH[v.lower()]
    except KeyError:
        raise argparse.ArgumentTypeError("Boolean value expected as string.") from None


def adjust_array_size(args):
    """nx grows to the next multiple of the block size, with a warning
    (commandline.py:22-26)."""
    b = args.blocksize
    grown = -(-args.spacesteps // b) * b
    if grown != args.spacesteps:
        args.spacesteps = grown
        warnings.warn(f"Adjusting array size to fit blocksize. New array size is {grown}.")


def euler_vortex(npx, npy, t=0.0, gamma=1.4):
    """Isentropic vortex sampled on the reference's CLI grid, endpoints
    included (the Euler plugin's createInitialConditions grid,
    equations/euler.py:33-56 with one rank)."""
    from pysweep_amd.equations import euler
    x = numpy.linspace(-5, 5, npx)
    y = numpy.linspace(-5, 5, npy)
    return numpy.stack(numpy.broadcast_arrays(*euler.analytical(x[:, None], y[None, :], t, gamma=gamma)))


def heat_mode(npx, npy, t=0.0, alpha=1.0):
    """sin(2 pi x) sin(2 pi y) exp(-8 pi^2 alpha t) on the reference's CLI
    grid (the heat plugin's createInitialConditions grid,
    equations/heat.py:66-93 with one rank)."""
    from pysweep_amd.equations import heat
    x = numpy.linspace(0, 1, npx)
    y = numpy.linspace(0, 1, npy)
    return heat.analyticalEquation(x[:, None], y[None, :], t, alpha)[None]


def _solver(args, ic, globals_, ops, its, cpu, gpu, output):
    from pysweep_amd.solver import Solver
    solver = Solver(initialConditions=ic.astype(numpy.float32).astype(numpy.float64), sendWarning=not args.ignore)
    solver.globals = globals_
    solver.dtypeStr = "float64"
    solver.dtype = numpy.dtype("float64")
    solver.verbose = args.verbose
    solver.operating = ops
    solver.intermediate = its
    solver.share = args.share
    solver.simulation = bool(args.swept)
    solver.blocksize = (args.blocksize, args.blocksize, 1)
    solver.setCPU(equations.path(cpu))
    solver.setGPU(equations.path(gpu))
    solver.exid = []
    solver.output = output
    solver.exact = bool(args.exact)
    solver.logfile = args.log
    solver.loadCPUModule()
    return solver


def run_euler(args):
    """utils/commandline.py:28-71."""
    if args.ignore:
        warnings.filterwarnings("ignore")
    adjust_array_size(args)
    gamma = 1.4
    dx = 10 / (args.spacesteps - 1)
    dt = 0.1 * dx
    ic = euler_vortex(args.spacesteps, args.spacesteps, gamma=gamma)
    solver = _solver(args, ic, [0, dt * args.timesteps, dt, dx, dx, gamma], 2, 2, "euler.py", "euler.cu",
                     "eulerOutput.hdf5")
    solver()
    return solver


def run_heat(args):
    """utils/commandline.py:73-115."""
    if args.ignore:
        warnings.filterwarnings("ignore")
    adjust_array_size(args)
    alpha = 1
    dx = 1 / (args.spacesteps - 1)
    dt = float(0.1 * dx ** 2 / alpha)
    ic = heat_mode(args.spacesteps, args.spacesteps, alpha=alpha)
    solver = _solver(args, ic, [0, dt * args.timesteps, dt, dx, dx, alpha, True], 1, 1, "heat.py", "heat.cu",
                     "heatOutput.hdf5")
    solver()
    return solver


FUNCTIONS = {"euler": run_euler, "heat": run_heat}


def parser():
    p = argparse.ArgumentParser(prog="pysweep_amd")
    p.add_argument("-f", "--function", nargs="?", default=None, choices=FUNCTIONS, type=str,
                   help="which example to run")
    p.add_argument("--swept", action="store_true", help="swept decomposition (default: standard)")
    p.add_argument("-b", "--blocksize", default=12, nargs="?", type=int,
                   help="block size b, a multiple of 2k (k points on each side of the operating point)")
    p.add_argument("-nx", "--spacesteps", default=24, nargs="?", type=int, help="points in x and y")
    p.add_argument("-nt", "--timesteps", default=10, nargs="?", type=int, help="time steps")
    p.add_argument("-s", "--share", default=1.0, nargs="?", type=float,
                   help="GPU share of the block rows, 0..1 (below 1 the host takes the rest; swept only)")
    p.add_argument("--clean", action="store_true", help="accepted for compatibility; no files are written")
    p.add_argument("--ignore", action="store_true", help="ignore warnings")
    p.add_argument("-v", "--verbose", action="store_true")
    p.add_argument("--exact", action="store_true", help="take exactly nt steps")
    p.add_argument("--log", default="log.yaml", help="run log (YAML) to append to; '' for none")
    return p


def command_line(argv=None):
    args = parser().parse_args(argv)
    if args.function is None:
        parser().print_usage()
        return None
    solver = FUNCTIONS[args.function](args)
    print(f"{'swept' if solver.simulation else 'standard'} {args.function} {args.spacesteps}^2 "
          f"b={args.blocksize} share={solver.share:g}: {solver.stages} stages in {solver.solve_seconds:.4f} s, "
          f"{solver.gcups:.3f} GCUPS")
    return solver


def main():
    command_line(sys.argv[1:])


if __name__ == "__main__":
    main()
