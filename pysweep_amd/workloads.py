"""Synthetic multi-problem workloads (seeded).

* ``batched_mix``: independent problems, half heat and half Euler, grid sides
  drawn uniformly from the multiples of 32 in [64, 1024] (non-square), heat
  initial conditions amp*sin(2 pi kx x) sin(2 pi ky y) with kx, ky in 1..4
  and amp ~ U(0.5, 2), Euler vortices with strength eps ~ U(1, 5) and centre
  uniform in [-5, 5]^2; 200 steps each, FP64.
* ``served_stream``: Poisson arrivals at a fixed rate of Euler vortex solves,
  sides log-uniform in [128, 2048] rounded to multiples of 32,
  eps ~ U(1, 5), step counts uniform in [100, 2000], 30% with FP32 state.

The isentropic vortex and the heat modes restate the reference's initial
conditions (pysweep/equations/euler.py:63-105, heat.py:90-93) on an
endpoint-free periodic grid.
"""
import numpy

from pysweep_amd.batch import Problem


def vortex_state(nx, ny, eps=5.0, xc=0.0, yc=0.0, gamma=1.4, L=5.0):
    """Isentropic vortex on [-L, L)^2 (conserved variables, shape (4,nx,ny))."""
    x = numpy.linspace(-L, L, nx, endpoint=False)
    y = numpy.linspace(-L, L, ny, endpoint=False)
    X, Y = numpy.meshgrid(x, y, indexing="ij")
    minf = numpy.sqrt(2 / gamma)
    beta = minf * eps * numpy.sqrt(2) / 4 / numpy.pi * numpy.exp(0.5)
    vinf = minf * numpy.sqrt(gamma)
    u0, v0 = vinf * numpy.cos(numpy.pi / 4), vinf * numpy.sin(numpy.pi / 4)
    dx0, dy0 = X - xc, Y - yc
    om = beta * numpy.exp(-0.5 * (dx0 ** 2 + dy0 ** 2))
    dT = -(gamma - 1) * om * om / 2
    rho = (1 + dT) ** (1 / (gamma - 1))
    u, v = u0 - dy0 * om, v0 + dx0 * om
    P = 1 / gamma * (1 + dT) ** (gamma / (gamma - 1))
    return numpy.stack([rho, rho * u, rho * v, P / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def heat_mode(nx, ny, kx=1, ky=1, amp=1.0):
    x = numpy.linspace(0, 1, nx, endpoint=False)
    y = numpy.linspace(0, 1, ny, endpoint=False)
    return (amp * numpy.outer(numpy.sin(2 * numpy.pi * kx * x), numpy.sin(2 * numpy.pi * ky * y)))[None]


def euler_problem(nx, ny, steps, eps=5.0, xc=0.0, yc=0.0, precision="fp64", b=None, tag=None):
    dx, dy = 10.0 / nx, 10.0 / ny
    dt = 0.1 * min(dx, dy)
    g = [0.0, steps * dt, dt, dx, dy, 1.4]
    return Problem("euler", vortex_state(nx, ny, eps, xc, yc), g, 2, precision=precision, b=b, tag=tag)


def heat_problem(nx, ny, steps, kx=1, ky=1, amp=1.0, precision="fp64", b=None, tag=None):
    dx, dy = 1.0 / nx, 1.0 / ny
    dt = 0.1 * dx * dx
    g = [0.0, steps * dt, dt, dx, dy, 1.0, True]
    return Problem("heat", heat_mode(nx, ny, kx, ky, amp), g, 1, precision=precision, b=b, tag=tag)


def batched_mix(n=512, steps=200, seed=0, lo=64, hi=1024):
    rng = numpy.random.default_rng(seed)
    sides = numpy.arange(lo, hi + 1, 32)
    out = []
    for i in range(n):
        nx, ny = int(rng.choice(sides)), int(rng.choice(sides))
        if i % 2 == 0:
            kx, ky = int(rng.integers(1, 5)), int(rng.integers(1, 5))
            out.append(heat_problem(nx, ny, steps, kx, ky, float(rng.uniform(0.5, 2.0)), tag=i))
        else:
            out.append(euler_problem(nx, ny, steps, float(rng.uniform(1, 5)), float(rng.uniform(-5, 5)),
                                     float(rng.uniform(-5, 5)), tag=i))
    return out


def served_stream(n=64, rate=4.0, seed=1, lo=128, hi=2048, steps=(100, 2000), fp32_share=0.3):
    """List of (arrival seconds, Problem)."""
    rng = numpy.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        side = lambda: int(max(lo, min(hi, round(float(numpy.exp(rng.uniform(numpy.log(lo), numpy.log(hi))))
                                                 / 32) * 32)))
        nx, ny = side(), side()
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        prec = "fp32" if rng.uniform() < fp32_share else "fp64"
        out.append((t, euler_problem(nx, ny, nsteps, float(rng.uniform(1, 5)), precision=prec, tag=i)))
    return out
