"""Command line (utils/commandline.py:28-144) and log.yaml (core/io.py:10-33)."""
import math
import os

import numpy as np
import pytest
import yaml

from pysweep_amd import cli


def test_parser_defaults_and_options():
    a = cli.parser().parse_args(["-f", "euler", "-nx", "100", "-nt", "7", "-b", "16", "--swept"])
    assert (a.function, a.spacesteps, a.timesteps, a.blocksize, a.swept) == ("euler", 100, 7, 16, True)
    d = cli.parser().parse_args([])
    assert (d.blocksize, d.spacesteps, d.timesteps, d.swept) == (12, 24, 10, False)
    with pytest.raises(SystemExit):
        cli.parser().parse_args(["-f", "sod"])


def test_str2bool():
    assert cli.str2bool("yes") and cli.str2bool("T") and not cli.str2bool("0")
    with pytest.raises(Exception):
        cli.str2bool("maybe")


def test_adjust_array_size_rounds_up_to_blocksize():
    a = cli.parser().parse_args(["-nx", "100", "-b", "16"])
    with pytest.warns(UserWarning):
        cli.adjust_array_size(a)
    assert a.spacesteps == 112
    a = cli.parser().parse_args(["-nx", "96", "-b", "16"])
    cli.adjust_array_size(a)
    assert a.spacesteps == 96


def test_initial_conditions_match_closed_forms():
    n = 33
    u = cli.heat_mode(n, n)
    x = np.linspace(0, 1, n)
    assert u.shape == (1, n, n)
    assert np.allclose(u[0], np.outer(np.sin(2 * np.pi * x), np.sin(2 * np.pi * x)), atol=1e-15)
    q = cli.euler_vortex(n, n)
    assert q.shape == (4, n, n)
    # far from the vortex the state is the free stream rho = 1, u = v = 1, p = 1/gamma
    g = 1.4
    assert abs(q[0, 0, 0] - 1) < 1e-6 and abs(q[1, 0, 0] - 1) < 1e-6 and abs(q[2, 0, 0] - 1) < 1e-6
    assert abs(q[3, 0, 0] - (1 / g / (g - 1) + 1.0)) < 1e-6
    # density and energy are radially symmetric about the centre
    assert np.allclose(q[0], q[0].T, atol=1e-14) and np.allclose(q[0], q[0][::-1, ::-1], atol=1e-14)
    # centre density (1 + dT)^(1/(gamma-1)) with dT = -(gamma-1) beta^2 / 2
    beta = math.sqrt(2 / g) * 5 * math.sqrt(2) / 4 / math.pi * math.exp(0.5)
    assert abs(q[0, n // 2, n // 2] - (1 - (g - 1) * beta * beta / 2) ** (1 / (g - 1))) < 1e-12


@pytest.mark.gpu
def test_cli_heat_runs_and_logs(device, tmp_path):
    """-f heat --swept --exact on 64^2: the oracle's 30 forward-Euler stages
    bitwise, and a log.yaml record with the run's figures."""
    from gpu_helpers import oracle_stages
    from oracle import heat as o_heat
    log = str(tmp_path / "log.yaml")
    s = cli.command_line(["-f", "heat", "-nx", "64", "-nt", "30", "-b", "8", "--swept", "--exact", "--log", log])
    n = 64
    dx = 1 / (n - 1)
    dt = float(0.1 * dx ** 2 / 1)
    c = o_heat.courant_py(1, dt, dx)
    ic = cli.heat_mode(n, n).astype(np.float32).astype(np.float64)
    want = oracle_stages(lambda q, qb, gg: o_heat.step_py(q[0], None, gg, c, True)[None], ic, 1, 30)
    assert np.array_equal(s.result, want)
    with open(log) as f:
        rec = yaml.safe_load(f)
    (entry,) = rec.values()
    assert entry["swept"] is True and entry["blocksize"] == 8 and entry["gcups"] > 0
    assert entry["array_shape"][1:] == [1, n, n]
