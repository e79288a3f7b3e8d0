"""GPU tests of instance-table launches (batching, continuous admission,
graph replay, FP32 state): every problem's result must not depend on what
else runs; FP64 results equal the CPU oracle bitwise."""
import numpy as np
import pytest

from gpu_helpers import normwise, o_euler, o_heat, oracle_stages

pytestmark = pytest.mark.gpu


def _euler_problem(nx, ny, steps, eps=3.0, xc=0.5, yc=-1.0, precision="fp64", b=None):
    from pysweep_amd.batch import Problem
    dx = 10.0 / nx
    dt = 0.1 * dx
    X, Y = np.meshgrid(np.linspace(-5, 5, nx, endpoint=False), np.linspace(-5, 5, ny, endpoint=False),
                       indexing="ij")
    ic = o_euler.vortex(X, Y, 0.0, 1.4, eps=eps, xc=xc, yc=yc)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    step = lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx)
    return Problem("euler", ic, g, 2, precision=precision, b=b), step


def _heat_problem(nx, ny, steps, fe=True, kx=1, ky=2, amp=1.3, precision="fp64", b=None):
    from pysweep_amd.batch import Problem
    dx = 1.0 / nx
    dt = 0.1 * dx * dx
    x = np.linspace(0, 1, nx, endpoint=False)
    y = np.linspace(0, 1, ny, endpoint=False)
    ic = (amp * np.outer(np.sin(2 * np.pi * kx * x), np.sin(2 * np.pi * ky * y)))[None]
    g = [0.0, steps * dt, dt, dx, dx, 1.0, fe]
    c = o_heat.courant_py(1.0, dt, dx)
    step = lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c, fe)[None]
    return Problem("heat", ic, g, 1 if fe else 2, precision=precision, b=b), step


def _mix():
    return [_euler_problem(64, 96, 9), _euler_problem(96, 64, 5, eps=4.0), _euler_problem(128, 128, 7),
            _heat_problem(64, 64, 30), _heat_problem(128, 64, 17, kx=2), _heat_problem(64, 128, 12, fe=False)]


def _run(problems, **kw):
    from pysweep_amd.batch import Engine
    eng = Engine(**kw)
    tickets = [eng.submit(p, arrival=0.0) for p in problems]
    eng.run()
    return tickets, eng


def test_batched_mix_matches_oracle_bitwise(device):
    mix = _mix()
    tickets, eng = _run([p for p, _ in mix])
    for t, (p, step) in zip(tickets, mix):
        steps = int((p.globals[1] - p.globals[0]) / p.globals[2])
        want = oracle_stages(step, p.ic, p.its, p.its * steps)
        assert np.array_equal(t.result, want), (p.plugin, p.ic.shape)
        assert t.conservation["drift"] < 1e-13, t.conservation


def test_alone_equals_in_mix(device):
    mix = _mix()
    tickets, _ = _run([p for p, _ in mix])
    for i in (0, 4):
        alone, _ = _run([mix[i][0]])
        assert np.array_equal(alone[0].result, tickets[i].result)


def test_fp32_alone_equals_in_mix(device):
    """FP32-state problems in a mix of FP64 and FP32 ones of both plugins,
    with swept and standard methods: each FP32 result equals the same
    problem's result run alone, bitwise."""
    mix = _mix() + [_euler_problem(96, 64, 7, eps=2.0, precision="fp32"),
                    _euler_problem(64, 64, 5, precision="fp32"),
                    _heat_problem(64, 128, 21, kx=3, precision="fp32"),
                    _heat_problem(128, 64, 15, kx=2, precision="fp32")]
    mix[-1][0].method = "standard"
    mix[-3][0].method = "standard"
    tickets, _ = _run([p for p, _ in mix])
    for i in range(len(mix) - 4, len(mix)):
        alone, _ = _run([mix[i][0]])
        r = alone[0].result  # FP32 state, returned as float64
        assert np.array_equal(r.astype(np.float32).astype(np.float64), r)
        assert np.array_equal(alone[0].result, tickets[i].result), i


def test_solo_constants_follow_the_problem_across_classes(device):
    """Two heat problems of different classes (b=8 and b=16) and different
    time steps whose solo runs share one standard module: the first runs
    part of its stages alone, the second runs to the end, then the first
    resumes.  The module's constants must be the first problem's again."""
    from pysweep_amd.batch import Engine, Problem
    p1, step1 = _heat_problem(64, 64, 40, b=8)
    p1.method = "standard"
    n, steps = 64, 30
    dx = 1.0 / n
    dt = 0.05 * dx * dx
    x = np.linspace(0, 1, n, endpoint=False)
    ic2 = np.outer(np.sin(2 * np.pi * 3 * x), np.sin(2 * np.pi * x))[None]
    p2 = Problem("heat", ic2, [0.0, steps * dt, dt, dx, dx, 1.0, True], 1, b=16, method="standard")
    c2 = o_heat.courant_py(1.0, dt, dx)
    step2 = lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c2, True)[None]
    a = Engine(max_rounds_ahead=5)
    t1 = a.submit(p1, arrival=0.0)
    a.step()
    assert 0 < t1.c < t1.c_end
    b = Engine()
    t2 = b.submit(p2, arrival=0.0)
    b.run()
    a.run()
    assert np.array_equal(t2.result, oracle_stages(step2, p2.ic, 1, steps))
    assert np.array_equal(t1.result, oracle_stages(step1, p1.ic, 1, 40))


@pytest.mark.parametrize("rounds", [1, 3])
def test_graph_replay_equals_eager(device, rounds):
    """Rounds replayed from captured hipGraphs of 1 or 3 rounds (remainders
    eager) give the same bits as eager rounds."""
    mix = _mix()
    a, _ = _run([p for p, _ in mix], use_graphs=True, graph_rounds=rounds)
    b, _ = _run([p for p, _ in mix], use_graphs=False)
    for x, y in zip(a, b):
        assert np.array_equal(x.result, y.result)


def test_continuous_admission_under_hbm_budget(device):
    """Staggered arrivals and a budget that fits about two problems: problems
    queue, are admitted as memory frees up, and still match the oracle."""
    from pysweep_amd.batch import Engine
    mix = _mix()
    eng = Engine(hbm_budget=3 * 8 * 4 * 128 * 128 * 8)
    tickets = [eng.submit(p, arrival=0.002 * i) for i, (p, _) in enumerate(mix)]
    eng.run()
    assert eng.peak_in_use <= eng.hbm_budget
    assert all(t.done for t in tickets)
    for t, (p, step) in zip(tickets, mix):
        steps = int((p.globals[1] - p.globals[0]) / p.globals[2])
        assert np.array_equal(t.result, oracle_stages(step, p.ic, p.its, p.its * steps))


@pytest.mark.parametrize("copy_stream", [True, False])
def test_reused_ring_after_async_retirement(device, copy_stream):
    """One problem at a time (max_active=1): each admitted problem takes the
    ring its predecessor just left, while that ring's final level may still
    be on its way to the host.  Stream order (the copy stream's, or the work
    stream's) must keep every result intact, and the pool must serve every
    problem from the first ring."""
    from pysweep_amd.batch import Engine, release_rings
    mix = [_heat_problem(64, 64, 20 + 3 * i, kx=1 + i % 3, amp=0.5 + 0.25 * i) for i in range(5)]
    release_rings()  # no free rings left by earlier engines
    eng = Engine(max_active=1, copy_stream=copy_stream)
    tickets = [eng.submit(p, arrival=0.0) for p, _ in mix]
    assert eng.run() is eng.finished
    assert [t.done for t in tickets] == [True] * 5
    assert sorted(map(id, eng.finished)) == sorted(map(id, tickets))  # each ticket finished once
    assert len(eng.rings.free) == 1 and not eng.running and eng.in_use == 0
    for t, (p, step) in zip(tickets, mix):
        steps = int(round((p.globals[1] - p.globals[0]) / p.globals[2]))
        assert np.array_equal(t.result, oracle_stages(step, p.ic, p.its, p.its * steps)), t.problem.globals


@pytest.mark.parametrize("copy_stream", [True, False])
def test_small_staging_ring_wraps_waits_and_grows(device, monkeypatch, copy_stream):
    """A staging ring far smaller than the traffic: admissions and
    retirements wait for earlier copies to free room (events of the work
    and copy streams, released in order), pieces wrap around the ring, and a
    level larger than the ring makes a bigger one; every result must still
    equal the oracle."""
    from pysweep_amd import batch
    monkeypatch.setattr(batch, "STAGING_BYTES", 1 << 16)  # 64 KiB
    monkeypatch.setattr(batch, "_FREE_RINGS", [])
    mix = _mix()  # levels of 32 KiB .. 512 KiB
    eng = batch.Engine(admit_bytes=1 << 16, copy_stream=copy_stream)
    assert eng._ring.size == 1 << 16
    tickets = [eng.submit(p, arrival=0.0) for p, _ in mix]
    eng.run()
    assert all(t.done for t in tickets)
    assert max(r.size for r in batch._FREE_RINGS) > 1 << 16  # grew for the largest level
    for t, (p, step) in zip(tickets, mix):
        steps = int(round((p.globals[1] - p.globals[0]) / p.globals[2]))
        assert np.array_equal(t.result, oracle_stages(step, p.ic, p.its, p.its * steps))


def test_async_copies_and_event_query(device):
    """Page-locked buffers, the copies that do not wait, and the event
    query that does not block."""
    from pysweep_amd import hip
    n = 1 << 20
    h, back = hip.PinnedArray(8 * n), hip.PinnedArray(8 * n + 64)
    d, d2 = hip.DeviceArray(8 * n), hip.DeviceArray(8 * n)
    want = np.arange(n, dtype=np.float64) * 0.5 - 7.0
    h.view(np.float64, (n,))[:] = want
    back.view(np.uint8, (8 * n + 64,))[:] = 0xAB
    h.upload_async(d.ptr, 8 * n)
    hip.dtod_async(d2.ptr, d.ptr, 8 * n)
    back.download_async(d2.ptr + 8, 8 * (n - 1), 64)  # offsets on both sides
    ev = hip.Event()
    ev.record()
    ev.synchronize()
    assert ev.query()
    got = back.view(np.uint8, (8 * n + 64,))
    assert (got[:64] == 0xAB).all() and (got[64 + 8 * (n - 1):] == 0xAB).all()
    assert np.array_equal(got[64:64 + 8 * (n - 1)].view(np.float64), want[1:])
    for a in (h, back, d, d2):
        a.close()


def test_block_size_choice_does_not_change_results(device):
    (p16, step), (p32, _) = _euler_problem(128, 96, 11, b=16), _euler_problem(128, 96, 11, b=32)
    (p8, _) = _euler_problem(128, 96, 11, b=8)
    r, _ = _run([p16, p32, p8])
    assert np.array_equal(r[0].result, r[1].result)
    assert np.array_equal(r[0].result, r[2].result)


def test_tail_only_problem(device):
    """Fewer stages than one pyramid: every stage runs on the standard kernel."""
    p, step = _euler_problem(64, 64, 1, b=32)
    r, _ = _run([p])
    assert np.array_equal(r[0].result, oracle_stages(step, p.ic, 2, 2))


def test_fp32_state_error_against_fp64_oracle(device):
    """FP32 state, FP64 arithmetic: L2 and Linf errors against the FP64
    oracle, stated per result; heat within 1e-6 relative."""
    (h32, hstep) = _heat_problem(128, 128, 40, precision="fp32")
    (e32, estep) = _euler_problem(128, 128, 20, precision="fp32")
    r, _ = _run([h32, e32])
    want_h = oracle_stages(hstep, h32.ic, 1, 40)
    want_e = oracle_stages(estep, e32.ic, 2, 40)
    lin_h = normwise(r[0].result, want_h)
    l2_h = float(np.linalg.norm(r[0].result - want_h) / np.linalg.norm(want_h))
    lin_e = normwise(r[1].result, want_e)
    l2_e = float(np.linalg.norm(r[1].result - want_e) / np.linalg.norm(want_e))
    print(f"fp32 heat: L2 {l2_h:.2e} Linf {lin_h:.2e}; fp32 euler: L2 {l2_e:.2e} Linf {lin_e:.2e}")
    assert lin_h <= 1e-6 and l2_h <= 1e-6
    assert np.isfinite(r[1].result).all()
    assert l2_e <= 1e-3
    # conserved sums in FP64 on the device: FP32 storage rounding moves them
    # by about one FP32 ulp per stage at most
    for t in r:
        c = t.conservation
        s0 = t.problem.ic.astype(np.float32).astype(np.float64).sum(axis=(1, 2))
        assert np.allclose(c["initial"], s0, rtol=1e-12)
        assert np.allclose(c["final"], t.result.sum(axis=(1, 2)), rtol=1e-12, atol=1e-9)
        print(t.problem.plugin, "fp32 conserved-sum drift", c["drift"])
        assert c["drift"] < 1e-5


def test_auto_method_and_round_batching_match_oracle(device):
    """method="auto" (the faster kernel per class: Euler takes the band
    method) and engines that run one round per host step (two while a
    band-method problem takes part) or many give the oracle's results
    bitwise."""
    mix = _mix()
    for p, _ in mix:
        p.method = "auto"
    a, eng = _run([p for p, _ in mix])
    b, _ = _run([p for p, _ in mix], max_rounds_ahead=1)
    for t, u, (p, step) in zip(a, b, mix):
        steps = int((p.globals[1] - p.globals[0]) / p.globals[2])
        want = oracle_stages(step, p.ic, p.its, p.its * steps)
        assert np.array_equal(t.result, want) and np.array_equal(u.result, want)
    assert {t.method for t in a} <= {"swept", "standard", "band"}


def test_tuner_records_the_fastest_candidate_and_results_do_not_depend_on_it(device, tmp_path):
    """Tuner.tune times each candidate (method, block size, LDS window)
    on a sample batch, stores the fastest, and every candidate gives the
    same bits."""
    import json

    from pysweep_amd.autotune import Tuner
    path = str(tmp_path / "tuning.json")
    tu = Tuner(path=path)

    def make():
        out = []
        for i, (nx, ny, kx) in enumerate(((64, 64, 1), (128, 64, 2), (64, 128, 3))):
            p, _ = _heat_problem(nx, ny, 24, kx=kx)
            p.tag = i
            out.append(p)
        return out

    cands = [{"b": 16}, {"b": 32}, {"b": 32, "slots": 4}, {"b": 32, "method": "standard"},
             {"b": 32, "method": "band"}]
    rec = tu.tune("heat", "fp64", 1, make, candidates=cands)
    assert rec["b"] in (16, 32) and rec["method"] in ("swept", "standard", "band")
    with open(path) as f:
        stored = json.load(f)["heat/fp64/its1"]
    assert stored["b"] == rec["b"] and len(stored["trials"]) == len(cands)
    first = tu.last_trials[0]["results"]
    for t in tu.last_trials[1:]:
        for tag, r in t["results"].items():
            assert np.array_equal(r, first[tag]), (t, tag)
    assert Tuner(path=path).method("heat", "fp64", 1) == rec["method"]


def test_tune_classes_feeds_the_default_engine(device, tmp_path):
    """Tuner.tune_classes tunes every class of a mix on a sample of its own
    problems, records the chosen occupancy, and an engine built on the
    stored records runs method="auto" problems to the oracle's bits."""
    from pysweep_amd.autotune import Tuner
    path = str(tmp_path / "tuning.json")
    mix = _mix()
    recs = Tuner(path=path).tune_classes([p for p, _ in mix], sample=3, steps=6)
    assert set(recs) == {"euler/fp64/its2", "heat/fp64/its1", "heat/fp64/its2"}
    for r in recs.values():
        assert r["method"] in ("swept", "standard", "band") and r["workgroups_per_cu"] >= 1
        assert len(r["trials"]) >= 3
    for p, _ in mix:
        p.method = "auto"
    tickets, eng = _run([p for p, _ in mix], tuner=Tuner(path=path))
    for t, (p, step) in zip(tickets, mix):
        rec = recs[f"{p.plugin}/fp64/its{p.its}"]
        assert t.method == rec["method"]
        steps = int((p.globals[1] - p.globals[0]) / p.globals[2])
        assert np.array_equal(t.result, oracle_stages(step, p.ic, p.its, p.its * steps))


@pytest.mark.parametrize("method", ["swept", "standard", "band"])
def test_solo_path_equals_batched_rounds(device, method):
    """A problem running alone uses the plain kernels (swept cycles, then the
    exact-step tail); the same problem through batched rounds, and the
    oracle, give the same bits."""
    p, step = _euler_problem(96, 64, 10, b=16)  # 20 stages: swept cycles and a tail
    p.method = method
    a, _ = _run([p])
    b, _ = _run([p], solo=False)
    want = oracle_stages(step, p.ic, 2, 20)
    assert np.array_equal(a[0].result, want) and np.array_equal(b[0].result, want)


@pytest.mark.parametrize("capacity", [1, 64 * 96 + 96 * 64, 1 << 19])
def test_oldest_first_policy_matches_oracle(device, capacity):
    """The "oldest_first" selection (rounds take problems in arrival order
    up to a cell capacity; capacity 1 runs them one at a time on the solo
    path, the middle one mixes solo and batched rounds) leaves every result
    bitwise the oracle's."""
    mix = _mix()
    from pysweep_amd.batch import Engine
    eng = Engine(policy="oldest_first", capacity=capacity)
    tickets = [eng.submit(p, arrival=0.01 * i) for i, (p, _) in enumerate(mix)]
    eng.run()
    for t, (p, step) in zip(tickets, mix):
        steps = int((p.globals[1] - p.globals[0]) / p.globals[2])
        want = oracle_stages(step, p.ic, p.its, p.its * steps)
        assert np.array_equal(t.result, want), (p.plugin, p.ic.shape, capacity)


def _band_mix():
    """Euler problems on the band method (grids narrower than one band of 56
    output columns, ragged last bands, non-square, odd step counts), an
    Euler problem of the same class on each of the other methods, and a heat
    problem beside them."""
    out = [_euler_problem(64, 96, 9), _euler_problem(96, 64, 5, eps=4.0), _euler_problem(128, 176, 7, eps=2.0),
           _euler_problem(160, 128, 6, xc=-1.0), _euler_problem(64, 64, 4), _euler_problem(96, 96, 8, eps=1.5),
           _euler_problem(48, 16, 5, eps=2.5), _euler_problem(16, 48, 3, yc=1.0)]
    for i in (0, 1, 2, 3, 6, 7):  # 16 and 48 columns: narrower than a band and than the wavefront
        out[i][0].method = "band"
    out[4][0].method = "standard"
    out[5][0].method = "swept"
    # heat on the band method: launches of 12 levels, the last one shorter
    # (30 forward-Euler stages: 12 + 12 + 6; RK2 17 steps: 12 + 12 + 10;
    # 5 stages: one short launch), beside a heat problem on the standard method
    heat = [_heat_problem(64, 64, 30), _heat_problem(64, 64, 30, kx=2, amp=0.8),
            _heat_problem(128, 64, 17, fe=False, ky=1), _heat_problem(64, 96, 5, kx=3)]
    for p, _ in heat[1:]:
        p.method = "band"
    return out + heat


def _check_oracle(tickets, mix, tag=""):
    for t, (p, step) in zip(tickets, mix):
        steps = int(round((p.globals[1] - p.globals[0]) / p.globals[2]))
        want = oracle_stages(step, p.ic, p.its, p.its * steps)
        assert np.array_equal(t.result, want), (p.plugin, p.method, p.ic.shape, tag)


@pytest.mark.parametrize("solo", [True, False])
def test_band_method_batched_matches_oracle_bitwise(device, solo):
    """Band-method problems share BandMulti launches with each other while
    the swept, standard and heat problems take their own kernels in the same
    rounds: every result is bitwise the oracle's, with the one-problem
    kernels taking over the last problem running (solo=True) or not."""
    mix = _band_mix()
    tickets, eng = _run([p for p, _ in mix], solo=solo)
    _check_oracle(tickets, mix, f"solo={solo}")
    assert all(t.conservation["drift"] < 1e-13 for t in tickets)


def test_band_method_hands_over_between_solo_and_batched_rounds(device):
    """A band-method problem runs alone on the one-problem kernels, then in
    batched rounds once others arrive, then alone again: each host step
    leaves it at the start of a step (the batched launch writes only a
    step's last level), and every result is bitwise the oracle's."""
    from pysweep_amd.batch import Engine
    mix = _band_mix()
    first, rest = mix[2], [mix[0], mix[4], mix[8]]
    eng = Engine(max_rounds_ahead=3)
    t0 = eng.submit(first[0], arrival=0.0)
    eng.step()
    assert t0.c > 1 and (t0.c - 1) % 2 == 0, t0.c  # 3 rounds asked, a whole number of steps run
    tickets = [eng.submit(p, arrival=0.0) for p, _ in rest]
    while not eng.idle():
        eng.step()
        if not t0.done:
            assert (t0.c - 1) % 2 == 0, t0.c
    eng.run()
    _check_oracle([t0] + tickets, [first] + rest)


def test_band_method_fp32_alone_equals_in_mix_and_standard(device):
    """FP32-state problems on the band method in a mix: each result equals
    the same problem run alone on the standard method, bitwise."""
    import dataclasses
    mix = _band_mix()
    f32 = [_euler_problem(96, 64, 7, eps=2.0, precision="fp32"), _euler_problem(64, 112, 5, precision="fp32")]
    for p, _ in f32:
        p.method = "band"
    tickets, _ = _run([p for p, _ in mix + f32])
    for i, (p, _) in enumerate(f32):
        alone, _ = _run([dataclasses.replace(p, method="standard")], solo=False)
        assert np.array_equal(alone[0].result, tickets[len(mix) + i].result), i


@pytest.mark.parametrize("rounds", [1, 3])
def test_band_method_graph_replay_equals_eager(device, rounds):
    """Band-method rounds replayed from captured hipGraphs of 1 or 3 rounds
    give the same bits as eager rounds."""
    mix = _band_mix()
    a, _ = _run([p for p, _ in mix], use_graphs=True, graph_rounds=rounds, solo=False)
    b, _ = _run([p for p, _ in mix], use_graphs=False, solo=False)
    for x, y in zip(a, b):
        assert np.array_equal(x.result, y.result)


def test_band_method_rejected_without_a_band_form(device, monkeypatch):
    """A class without an instance-table band form: method="band" is
    refused at admission with a clear error."""
    from pysweep_amd import batch, kernels
    from pysweep_amd.batch import Engine
    monkeypatch.setattr(kernels, "band_multi_capable", lambda *a: False)
    p, _ = _heat_problem(64, 64, 4, b=8)  # a class of its own (b=8)
    p.method = "band"
    eng = Engine(tuner=None)
    eng.submit(p, arrival=0.0)
    with pytest.raises(ValueError, match="band sweep"):
        eng.run()


def test_engines_share_free_rings(device):
    """A finished engine leaves its rings to the next engine of the same
    context (no device free between them), within the new engine's HBM
    budget; release_rings frees them."""
    from pysweep_amd.batch import Engine, release_rings
    release_rings()
    mix = _mix()
    a, ea = _run([p for p, _ in mix])
    held = {id(r) for r in ea.rings.free}
    assert ea.rings.bytes > 0 and ea.in_use == 0
    b, eb = _run([p for p, _ in mix])
    assert eb.rings is ea.rings and {id(r) for r in eb.rings.free} == held  # every ring reused, none new
    for x, y in zip(a, b):
        assert np.array_equal(x.result, y.result)
    total = ea.rings.bytes
    small = Engine(hbm_budget=total // 2)
    assert 0 < total and small.rings.bytes <= total // 2
    assert Engine(copy_stream=False).rings is not ea.rings  # copies queued elsewhere: rings of their own
    release_rings()
    assert Engine().rings.bytes == 0


def test_engine_on_a_context_of_its_own(device):
    """An engine on a thread's own context: its copies go through a copy
    context made on that context's device, results equal the process
    context's engine bitwise, and closing the context releases its pooled
    rings and its copy context."""
    from pysweep_amd import batch, hip
    mix = _mix()
    want, _ = _run([p for p, _ in mix])
    ctx = hip.Context(hip.current_device())
    with ctx:
        got, eng = _run([p for p, _ in mix])
        key = hip.current_context()
        assert key in batch._COPY_CONTEXTS and (key, True) in batch._DEVICE_RINGS
    for x, y in zip(want, got):
        assert np.array_equal(x.result, y.result)
    ctx.close()
    assert key not in batch._COPY_CONTEXTS and (key, True) not in batch._DEVICE_RINGS


@pytest.mark.parametrize("policy,capacity", [("short_first", 1 << 13), ("short_first", 1 << 15),
                                             ("oldest_first", 1 << 14)])
def test_band_method_under_selection_policies(device, policy, capacity):
    """Round selection that leaves problems out of some rounds and takes
    them back later (least remaining work first, or arrival order, up to a
    small cell capacity), with staggered arrivals: band-method problems sit
    out and resume at the start of a step, and every result is bitwise the
    oracle's."""
    from pysweep_amd.batch import Engine
    mix = _band_mix()
    eng = Engine(policy=policy, capacity=capacity, poll_rounds=3)
    tickets = [eng.submit(p, arrival=0.003 * i) for i, (p, _) in enumerate(mix)]
    eng.run()
    _check_oracle(tickets, mix, f"{policy} {capacity}")


def test_heat_band_method_hands_over_with_a_short_last_launch(device):
    """A heat problem on the band method (launches of 12 stages) runs alone,
    then beside another, then alone again; host steps of whole launches may
    ask for more rounds than its stages left, and the one-problem path stops
    at its last stage: results bitwise the oracle's."""
    from pysweep_amd.batch import Engine
    a, sa = _heat_problem(64, 64, 41, kx=2)  # 41 stages: 3 launches of 12 and one of 5
    b, sb = _heat_problem(64, 128, 20, fe=False)  # RK2: 40 stages, 12 + 12 + 12 + 4
    a.method = b.method = "band"
    eng = Engine(max_rounds_ahead=5)
    ta = eng.submit(a, arrival=0.0)
    eng.step()
    assert ta.c == 13, ta.c  # 5 rounds asked, one whole launch run
    tb = eng.submit(b, arrival=0.0)
    eng.run()
    _check_oracle([ta, tb], [(a, sa), (b, sb)])


def test_heat_band_method_fp32_equals_standard(device):
    """FP32-state heat problems on the band method in a mix: each result
    equals the same problem run alone on the standard method, bitwise."""
    import dataclasses
    mix = _band_mix()
    f32 = [_heat_problem(64, 64, 25, kx=2, precision="fp32"), _heat_problem(128, 64, 14, precision="fp32")]
    for p, _ in f32:
        p.method = "band"
    tickets, _ = _run([p for p, _ in mix + f32])
    for i, (p, _) in enumerate(f32):
        alone, _ = _run([dataclasses.replace(p, method="standard")], solo=False)
        assert np.array_equal(alone[0].result, tickets[len(mix) + i].result), i
