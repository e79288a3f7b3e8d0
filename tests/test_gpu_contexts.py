"""Contexts of their own (ps_context_*): solves driven from distinct threads,
each on its own context and work stream, give the results of the same solves
run one after another on the process context, bitwise, and the oracle's."""
import threading

import numpy as np
import pytest

from gpu_helpers import euler_case, gpu_solve, heat_case, normwise, oracle_stages

pytestmark = pytest.mark.gpu


def _cases():
    # two heat problems on the same kernel module spec (same n, b, s) with
    # different time steps: their __constant__ globals differ, so each
    # context must hold its own copy of the module
    h1, g1, st1 = heat_case(128, 1, 40)
    h2, g2, st2 = heat_case(128, 1, 40, seed=3)
    g2 = list(g2)
    g2[2] *= 0.5  # dt
    g2[1] = 40 * g2[2]
    from oracle import heat as o_heat
    c2 = o_heat.courant_py(1.0, g2[2], g2[3])
    st2 = (lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c2, 1)[None])
    e1, ge, ste = euler_case(64, 12)
    return [("heat", h1, g1, 1, 1, 8, st1, 40), ("heat", h2, g2, 1, 1, 8, st2, 40),
            ("euler", e1, ge, 2, 2, 16, ste, 24)]


def test_solves_on_separate_contexts_and_threads_match_solo(device):
    from pysweep_amd import hip
    cases = _cases()
    solo = [gpu_solve(name, ic, g, k, s, b).final_state.copy() for name, ic, g, k, s, b, _, _ in cases]
    ctxs = [hip.Context(0) for _ in cases]
    out, errors = [None] * len(cases), []
    barrier = threading.Barrier(len(cases))

    def work(i):
        name, ic, g, k, s, b, _, _ = cases[i]
        try:
            with ctxs[i]:
                assert hip.current_context() == ctxs[i].handle
                barrier.wait(timeout=60)
                # three solves in a row per thread, so the threads overlap on the GPU
                for _ in range(3):
                    out[i] = gpu_solve(name, ic, g, k, s, b).final_state.copy()
                hip.synchronize()
        except Exception as e:  # reported in the main thread
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    for c in ctxs:
        c.close()
    assert not errors, errors
    assert hip.current_context() is None
    for i, (name, ic, g, k, s, b, step, stages) in enumerate(cases):
        assert np.array_equal(out[i], solo[i]), (name, i)
        want = oracle_stages(step, ic, s, stages)
        assert normwise(out[i], want) <= 1e-12, (name, i)
    # the two heat problems really had different constants
    assert not np.array_equal(out[0], out[1])


def test_context_misuse_is_reported(device):
    from pysweep_amd import hip
    ctx = hip.Context(0)
    with ctx:
        with pytest.raises(hip.HipError, match="current on the calling thread"):
            ctx.close()
    ctx.close()
    with pytest.raises(hip.HipError, match="no context is current"):
        hip.check(hip._ctx_pop(), "ps_context_pop")
