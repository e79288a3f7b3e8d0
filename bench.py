"""Benchmark: 2-D Euler GCUPS, swept vs standard, on one MI355X.

Workload (the reference's GPU configuration): isentropic vortex (eps=5 at the
origin, free stream (1,1)), periodic [-5,5]^2, nx = ny = 1024, gamma = 1.4,
dt = 0.1*dx, blocksize 16, operating_points 2, intermediate_steps 2,
500 full steps (1000 RK2 stages), float64, state resident in HBM.

One bench step = one complete solve of that workload.  ``value`` is the
swept solve's GCUPS = nx*ny*500 / (seconds per solve) / 1e9, the swept solve
being the product's default for the Euler plugin: the band kernel (one RK2
step per launch, every level in registers, the blocks' sloped edges
recomputed instead of bridged).  The standard kernel's GCUPS and the five
phase kernels' (the swept rule's pyramids, bridges and octahedra, b=16) on
the same inputs are reported beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line (rank 0).  ``roofline`` is for the dominant swept
kernel (the largest total launch time in the warm-up solve), with its
average launch duration measured in the timed region by HIP events on the
work stream around one launch in four of that kernel; ``traffic`` comes
from the committed rocprofv3 counter summary when present.
``cpu_baseline`` times the numpy oracle (the reference's numpy step order)
on a bounded sample of the same grid.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Euler 2D GCUPS swept vs standard, 1 GPU; served-solve p50/p99 latency"
FLOPS_PER_CELL_STAGE = 626          # euler.cu:22-244 counted per stage (1252 per RK2 step)
FP64_PEAK_TFLOPS = 78.6             # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_euler1024_b16.json")
#: the kernel build the committed counter file was measured on (bumped when
#: the swept or standard kernels change)
KERNEL_BUILD = 5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--b", type=int, default=16)
    ap.add_argument("--nsteps", type=int, default=500)
    ap.add_argument("--cpu-stages", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["swept", "standard", "phases", "both", "all"], default="all",
                    help="legs of the headline: swept (band kernel), standard, phases (five phase kernels); "
                         "both = swept and standard, all = the three")
    ap.add_argument("--batch", type=int, default=512, help="problems of the batched workload (0: skip)")
    ap.add_argument("--batch-legs", default="tuned,swept,standard,band,plain",
                    help="runs of the batched workload: engine tuned / with swept, standard or band kernels, "
                         "plain one at a time")
    ap.add_argument("--serve", type=int, default=48, help="requests of the served stream (0: skip)")
    ap.add_argument("--rate", type=float, default=4.0, help="served-stream arrival rate per second")
    ap.add_argument("--large", type=int, default=1, help="large single instances (heat 8192^2, Euler 2048^2)")
    ap.add_argument("--large-cases", default="heat,euler", help="large instances to run")
    ap.add_argument("--large-reals", default="double,float", help="state types of the large instances")
    ap.add_argument("--large-steps", type=int, default=1000)
    ap.add_argument("--precision-study", type=int, default=1,
                    help="FP32 error beside the FP64 one-ulp sensitivity (with --large)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's copy kernel on 32 MiB (three times): the reference transfer "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--generic", action="store_true", help="per-cell plugin step instead of the level update")
    ap.add_argument("--slots", type=int, default=None)
    ap.add_argument("--threads", type=int, default=None)
    return ap.parse_args()


def setup(n, b, nsteps, generic=False, slots=None, threads=None):
    """The headline's runs: the default swept solve (the band kernel; the
    phase kernels with --generic, whose per-cell step the band sweep does not
    take), the five phase kernels, and the standard kernel."""
    import numpy as np

    from pysweep_amd import device, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path

    hip.init(0)
    dx = 10.0 / n
    dt = 0.1 * dx
    ic = workloads.vortex_state(n, n)
    g = [0.0, nsteps * dt, dt, dx, dx, 1.4]
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    ph = device.SweptRun(eq_path("euler.cu"), 4, 2, 2, n, n, b, generic=generic, slots=slots, threads=threads)
    sw = ph if generic else device.BandRun(eq_path("euler.cu"), 4, 2, 2, n, n)
    st = device.StandardRun(eq_path("euler.cu"), 4, 2, 2, n, n, generic=generic)
    for r in {id(x): x for x in (sw, ph, st)}.values():
        r.set_globals(cpu, g)
        r.keep_ic_on_device(ic)
    times = schedule.swept_times(nsteps, b, 2, 2, exact=True)
    return np, hip, ic, g, sw, ph, st, times


SAMPLE_EVERY = 4  # timed region: HIP events around one launch in 4 of the dominant kernel


def time_solves(hip, run_once, reset, K, W):
    """Seconds per solve over K timed solves after W warm-ups.  The warm-ups
    time every launch (the per-kernel breakdown, and the dominant kernel: the
    largest total launch time); the timed solves put HIP events around one
    launch in SAMPLE_EVERY of the dominant kernel only, so the events add
    little to the time they measure.  Returns (seconds, timed-region stats,
    warm-up stats)."""
    warm = hip.RunStats()
    for _ in range(W):
        reset()
        run_once(warm)
    hip.synchronize()
    d = warm.as_dict()
    timed = [k for k in d if d[k]["launches"]]
    if timed:
        stats = hip.RunStats.sampled([max(timed, key=lambda k: d[k]["ms"])], SAMPLE_EVERY)
    else:
        stats = hip.RunStats()
    t0 = time.perf_counter()
    for _ in range(K):
        reset()
        run_once(stats)
    hip.synchronize()
    return (time.perf_counter() - t0) / K, stats, warm


def _cpu_strip(args):
    """Worker of cpu_baseline: `stages` RK2 stages of the numpy oracle on an
    x-strip extended by k*stages rows on each side (the rows a stage cannot
    compute correctly shrink by k per stage, so the strip's own rows are
    exact); returns the strip's rows of the final level and the seconds."""
    q, qb, lo, hi, halo, stages, dt, dx = args
    from oracle import euler as o_euler
    t0 = time.perf_counter()
    for gts in range(1, stages + 1):
        nq = o_euler.step_py(q, qb, gts, 1.4, dt, dx, dx)
        qb, q = q, nq
    return q[:, halo:halo + (hi - lo)], time.perf_counter() - t0


#: host cores a GPU may use on this machine (its share of the host's CPUs)
HOST_CORES_PER_GPU = 16


def host_cores():
    """Cores this process may run on, at most the GPU's share of the host."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(HOST_CORES_PER_GPU, n))


def cpu_baseline(ic, g, stages, workers=None):
    """numpy oracle (the reference's numpy step order) on the full grid for
    a few stages, x-strips over `workers` processes (every core the GPU's
    share of the host gives: host_cores()); the result is checked against a
    one-process run of the first stages on the same grid."""
    import multiprocessing as mp

    import numpy as np

    from oracle import euler as o_euler
    dt, dx = g[2], g[3]
    workers = workers or host_cores()
    nv, nx, ny = ic.shape
    halo = 2 * stages
    bounds = [(nx * i // workers, nx * (i + 1) // workers) for i in range(workers)]
    jobs = []
    for lo, hi in bounds:
        rows = np.arange(lo - halo, hi + halo) % nx
        jobs.append((ic[:, rows], ic[:, rows], lo, hi, halo, stages, dt, dx))
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        pool.map(_cpu_strip, [(j[0][:, :2 * halo + 1], j[1][:, :2 * halo + 1]) + j[2:] for j in jobs[:workers]])
        t0 = time.perf_counter()
        parts = pool.map(_cpu_strip, jobs)
        sec = time.perf_counter() - t0
    final = np.concatenate([p[0] for p in parts], axis=1)
    # check: the strips agree with the whole-grid oracle
    q, qb = ic.copy(), ic.copy()
    for gts in range(1, stages + 1):
        nq = o_euler.step_py(q, qb, gts, 1.4, dt, dx, dx)
        qb, q = q, nq
    n = nx * ny
    steps = stages / 2
    return {"value": n * steps / sec / 1e9, "unit": "GCUPS", "cores": workers, "kind": "port",
            "host_cpus_visible": os.cpu_count(), "cores_note": f"the GPU's share of the host: at most "
                                                              f"{HOST_CORES_PER_GPU} cores per GPU",
            "sample": f"{stages} RK2 stages ({steps:g} steps) of the {nx}x{ny} Euler vortex with the vectorized "
                      f"numpy oracle in the reference's numpy step order, x-strips over {workers} processes, "
                      f"{sec:.2f} s",
            "seconds": sec, "matches_whole_grid_oracle": bool(np.array_equal(final, q)),
            "np_version": np.__version__}


def _size_bucket(p):
    cells = p.ic.shape[1] * p.ic.shape[2]
    return "small(<=256^2)" if cells <= 256 * 256 else ("medium(<=640^2)" if cells <= 640 * 640 else "large")


def batched_workload(n, seed=0, legs=("tuned", "swept", "standard", "band", "plain"), buckets=True):
    """Independent heat and Euler problems, 200 steps each, FP64: the batched
    engine (instance-table launches, graph-replayed rounds) against solving
    them one at a time with the plain standard kernel.

    The "tuned" engine is the default one after ``Tuner.tune_classes`` has
    picked each class's block size, method and workgroup shapes on a sample
    of its problems (40 steps each; the time it took is reported); the
    "swept" and "standard" engines force one method with the default
    shapes, the "band" engine the band method (both plugins).  ``buckets`` adds a per-size breakdown: each size bucket run
    alone on the tuned engine and one at a time.  ``legs`` narrows the study
    (a profile of one engine run uses it)."""
    from pysweep_amd import autotune, baseline, batch, hip, workloads
    probs = workloads.batched_mix(n=n, steps=200, seed=seed)
    cells = sum(p.ic.shape[1] * p.ic.shape[2] * 200 for p in probs)
    out = {"problems": n, "steps": 200, "cell_steps": cells}
    results = {}
    tuner = None
    if "tuned" in legs:
        tuner = autotune.Tuner(path=None)
        t0 = time.perf_counter()
        recs = tuner.tune_classes(probs, sample=24, steps=40, save=False)
        out["tuning"] = {"seconds": time.perf_counter() - t0, "sample_problems_per_class": 24, "sample_steps": 40,
                         "choices": {k: {kk: v for kk, v in r.items() if kk != "trials"} for k, r in recs.items()},
                         "trials": {k: [{kk: v for kk, v in t.items() if kk in ("b", "method", "slots", "stile",
                                                                                "sthreads", "gcups",
                                                                                "workgroups_per_cu")}
                                        for t in r["trials"]] for k, r in recs.items()}}

    def engine_run(ps, method, tuner_):
        for p in ps:
            p.method = method(p) if callable(method) else method
        batch.release_rings()  # each engine allocates its own rings (its peak HBM is its own)
        eng = batch.Engine(tuner=tuner_)
        for p in ps:
            eng.submit(p, arrival=0.0)
        hip.synchronize()
        t0 = time.perf_counter()
        eng.run()
        hip.synchronize()
        return time.perf_counter() - t0, eng

    for label, method, tu in (("engine_tuned", "auto", tuner), ("engine_swept", "swept", None),
                              ("engine_standard", "standard", None), ("engine_band", "band", None)):
        if label.split("_")[1] not in legs:
            continue
        sec, eng = engine_run(probs, method, tu)
        results[label] = {t.problem.tag: t.result for t in eng.finished}
        drift = max(t.conservation["drift"] for t in eng.finished)
        out[label] = {"seconds": sec, "solves_per_s": n / sec, "gcups": cells / sec / 1e9, "rounds": eng.rounds,
                      "peak_hbm_gb": eng.peak_in_use / 1e9, "max_conserved_sum_drift": drift}
    if "plain" not in legs:
        return out
    per = {}
    t0 = time.perf_counter()
    same = 0
    for p in probs:
        t1 = time.perf_counter()
        r = baseline.solve_plain(p)
        per[p.tag] = time.perf_counter() - t1
        same += int(all(numpy_equal(r, res[p.tag]) for res in results.values()))
    sec = time.perf_counter() - t0
    out["one_at_a_time_plain"] = {"seconds": sec, "solves_per_s": n / sec, "gcups": cells / sec / 1e9}
    out["bitwise_equal_results"] = same
    if results:
        key = "engine_tuned" if "engine_tuned" in out else max(results, key=lambda k: out[k]["solves_per_s"])
        out["engine_over_plain"] = out[key]["solves_per_s"] / out["one_at_a_time_plain"]["solves_per_s"]
        out["engine_over_plain_leg"] = key
    if buckets and tuner is not None:
        rows = {}
        for name in sorted({_size_bucket(p) for p in probs}):
            ps = [p for p in probs if _size_bucket(p) == name]
            esec, _ = engine_run(ps, "auto", tuner)
            psec = sum(per[p.tag] for p in ps)
            rows[name] = {"problems": len(ps), "cell_steps": sum(p.ic.shape[1] * p.ic.shape[2] * 200 for p in ps),
                          "engine_s": esec, "plain_s": psec, "engine_over_plain": psec / esec}
        out["size_buckets"] = rows
    return out


HEAT_FLOPS_PER_CELL_STAGE = 12      # heat.cu:11-24 as written


def large_instances(names=("heat", "euler"), steps=1000, reals=("double", "float")):
    """Large single instances, 1000 steps, swept vs standard, FP64 and FP32
    state: heat 8192^2 (phase kernels with b=32 and b=64, and the band
    kernel) and Euler 2048^2 (phase kernels with b=32 and b=16, and the band
    kernel).  FP32
    errors are stated against the FP64 result (itself bitwise equal to the
    CPU oracle, which the parity tests check at smaller sizes).  ``names``,
    ``steps`` and ``reals`` narrow the study (the profiling runs use them)."""
    import numpy as np

    from pysweep_amd import device, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    out = []
    cases = []
    n = 8192
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    heat_g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    cases.append(("heat", n, (32, 64), 1, 1, 1, heat_g, ("heat.cu", "heat_fp32.cu", "heat.py")))
    n = 2048
    dx = 10.0 / n
    dt = 0.1 * dx
    cases.append(("euler", n, (32, 16), 4, 2, 2, [0.0, steps * dt, dt, dx, dx, 1.4],
                  ("euler.cu", "euler_fp32.cu", "euler.py")))
    for name, n, bs, nv, k, s, g, (c64, c32, py) in cases:
        if name not in names:
            continue
        ic = workloads.heat_mode(n, n) if name == "heat" else workloads.vortex_state(n, n)
        cpu = plugins.load_cpu_module(eq_path(py))
        cpu.set_globals(*g)
        finals = {}
        for real, cu in (("double", c64), ("float", c32)):
            if real not in reals:
                continue
            runs = [("standard", None)] + [("swept", b) for b in bs]
            # the swept solve of the band kernel (levels in registers,
            # halos recomputed instead of bridged)
            runs.append(("band", None))
            for method, b in runs:
                if method == "band":
                    r = device.BandRun(eq_path(cu), nv, k, s, n, n, real=real)
                    times = schedule.standard_times(steps, 8, k, s)
                    go = (lambda st, r=r, times=times: r.run(times, stats=st))
                    last = s * steps + s - 1
                elif method == "swept":
                    r = device.SweptRun(eq_path(cu), nv, k, s, n, n, b, real=real)
                    times = schedule.swept_times(steps, b, k, s, exact=True)
                    go = (lambda st, r=r, times=times: r.run(times, stats=st))
                    last = times.stages + times.tail + s - 1
                else:
                    r = device.StandardRun(eq_path(cu), nv, k, s, n, n, real=real)
                    go = (lambda st, r=r: r.run(1, s * steps + 1, stats=st))
                    last = s * steps + s - 1
                r.set_globals(cpu, g)
                r.keep_ic_on_device(ic)
                r.reset()
                warm = hip.RunStats()
                go(warm)  # warm-up, every launch timed: the dominant kernel
                r.reset()
                hip.synchronize()
                dw = warm.as_dict()
                kern = max(dw, key=lambda kk: dw[kk]["ms"])
                st = hip.RunStats.sampled([kern], SAMPLE_EVERY)
                t0 = time.perf_counter()
                go(st)
                hip.synchronize()
                sec = time.perf_counter() - t0
                finals[(real, method, b)] = r.level(last)
                d = st.as_dict()
                rb = 8 if real == "double" else 4
                if kern == "Standard":
                    cs = n * n
                elif kern == "Band":
                    cs = n * n * r.spec.levels
                else:
                    cs = (n // b) ** 2 * schedule.phase_cell_stages(kern, b, k, b // (2 * k) - 1)
                flops = FLOPS_PER_CELL_STAGE if name == "euler" else HEAT_FLOPS_PER_CELL_STAGE
                avg = d[kern]["avg_ms"] / 1e3
                rec = {"case": f"{name}-{n}x{n}", "method": method, "b": b, "state": "fp64" if rb == 8 else "fp32",
                       "gcups": n * n * steps / sec / 1e9, "seconds": sec, "dominant_kernel": kern,
                       "dominant_avg_launch_ms": d[kern]["avg_ms"],
                       "dominant_fp64_tflops": cs * flops / avg / 1e12,
                       "dominant_fp64_frac": cs * flops / avg / 1e12 / FP64_PEAK_TFLOPS,
                       # compulsory HBM bytes per step: heat 16 (read + write), Euler 160
                       # (stage 1: 32 R + 32 W; stage 2: 64 R + 32 W), halved for FP32 state
                       "algorithmic_gbs": n * n * steps * (16 if name == "heat" else 160) * rb / 8 / sec / 1e9}
                rec["algorithmic_hbm_frac"] = rec["algorithmic_gbs"] / HBM_PEAK_GBS
                if method == "band":
                    # per launch: one read and one write of every cell
                    # (levels stages), the recomputed halo columns aside
                    rec["band_levels"], rec["band_rows"] = r.spec.levels, r.rows
                    rec["band_launch_hbm_gbs"] = n * n * 2 * rb / avg / 1e9
                out.append(rec)
                r.close()
        ref = finals.get(("double", "standard", None))
        for key, f in finals.items() if ref is not None else ():
            if key[0] == "double":
                if not np.array_equal(f, ref):
                    out.append({"case": f"{name}-{n}x{n}", "mismatch": list(map(str, key))})
                continue
            e = f - ref
            for rec in out:
                if rec.get("case") == f"{name}-{n}x{n}" and rec.get("state") == "fp32" and \
                        (rec["method"], rec["b"]) == key[1:]:
                    rec["fp32_rel_L2"] = float(np.sqrt((e * e).sum() / (ref * ref).sum()))
                    rec["fp32_rel_Linf"] = float(np.abs(e).max() / np.abs(ref).max())
        del finals
    return out


def precision_sensitivity(n=1024, b=16, steps=(10, 100, 1000)):
    """Euler vortex n^2 (swept b): the FP32-state solve's error against the
    FP64 solve, beside the FP64 solve's own change when every value of the
    initial condition moves by one ulp.  The two grow alike: the limiter
    amplifies any perturbation, however small, to the same level."""
    import numpy as np

    from pysweep_amd import device, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    ic = workloads.vortex_state(n, n)
    ulp = np.nextafter(ic, np.where(ic >= 0, np.inf, -np.inf))
    out = []
    for st in steps:
        dx = 10.0 / n
        dt = 0.1 * dx
        g = [0.0, st * dt, dt, dx, dx, 1.4]
        cpu = plugins.load_cpu_module(eq_path("euler.py"))
        cpu.set_globals(*g)
        times = schedule.swept_times(st, b, 2, 2, exact=True)
        res = {}
        for name, cu, real, init in (("fp64", "euler.cu", "double", ic), ("fp32", "euler_fp32.cu", "float", ic),
                                     ("fp64_ulp", "euler.cu", "double", ulp)):
            r = device.SweptRun(eq_path(cu), 4, 2, 2, n, n, b, real=real)
            r.set_globals(cpu, g)
            r.upload_ic(init)
            last = r.run(times)
            hip.synchronize()
            res[name] = r.level(last)
            r.close()
        ref = res["fp64"]
        rec = {"case": f"euler-{n}x{n}-b{b}", "steps": st}
        for k in ("fp32", "fp64_ulp"):
            e = res[k] - ref
            rec[k + "_rel_L2"] = float(np.sqrt((e * e).sum() / (ref * ref).sum()))
            rec[k + "_rel_Linf"] = float(np.abs(e).max() / np.abs(ref).max())
        out.append(rec)
    return out


def product_host_path(n=1024, b=16, steps=8):
    """The product's own host path on a prefix of the headline workload: the
    Euler vortex at n^2 with share = 0, i.e. the swept decomposition computed
    on the host, level by level with the plugin's Python step over each
    phase's cells (the reference's CPU path, core/process.py:81-111 and
    core/geometry.py:26-36), one process.  ``steps`` steps with the
    reference's step quantization (b=16: P=3; 8 steps give G=4 and 15 stages)."""
    from pysweep_amd import Solver, workloads
    from pysweep_amd.equations import path as eq_path
    dx = 10.0 / n
    dt = 0.1 * dx
    ic = workloads.vortex_state(n, n)
    sol = Solver(ic, sendWarning=False)
    sol.simulation, sol.operating, sol.intermediate, sol.blocksize = True, 2, 2, (b, b, 1)
    sol.share = 0.0
    sol.exact = False
    sol.globals = [0.0, steps * dt, dt, dx, dx, 1.4]
    sol.setCPU(eq_path("euler.py"))
    sol.setGPU(eq_path("euler.cu"))
    sol.loadCPUModule()
    sol()
    cell_steps = n * n * sol.stages / 2
    return {"value": cell_steps / sol.solve_seconds / 1e9, "unit": "GCUPS", "cores": 1, "kind": "port",
            "sample": f"share=0 swept solve of the {n}x{n} Euler vortex on the host (the product's CPU path: "
                      f"the plugin's Python step over each phase level's cells), b={b}, {sol.stages} RK2 stages "
                      f"({sol.stages // 2} steps), one process, {sol.solve_seconds:.2f} s",
            "seconds": sol.solve_seconds}


def heat_cpu_case():
    """The reference's CPU-runnable case: heat 384^2, b=8, forward Euler,
    100 steps (99 stages with the reference's quantization), swept, from
    sin(2 pi x) sin(2 pi y) plus seeded noise: on the GPU alone (share 1),
    on the host alone (share 0: the plugin's Python step over each level's
    cells, one core), and split (share 0.5); all three bitwise equal."""
    import numpy as np

    from pysweep_amd import Solver
    from pysweep_amd.equations import path as eq_path
    from pysweep_amd import workloads
    n = 384
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    rng = np.random.default_rng(0)
    ic = workloads.heat_mode(n, n) + 1e-3 * rng.uniform(-1, 1, (1, n, n))
    out = {"workload": "heat-384x384-b8-100steps-fe-swept"}
    finals = {}
    for share in (1.0, 0.5, 0.0):
        sol = Solver(ic, sendWarning=False)
        sol.simulation, sol.operating, sol.intermediate, sol.blocksize = True, 1, 1, (8, 8, 1)
        sol.share = share
        sol.globals = [0.0, 100 * dt, dt, dx, dx, 1.0, True]
        sol.setCPU(eq_path("heat.py"))
        sol.setGPU(eq_path("heat.cu"))
        sol.loadCPUModule()
        sol()
        finals[share] = sol.final_state
        out[f"share_{share:g}"] = {"gcups": sol.gcups, "seconds": sol.solve_seconds, "stages": sol.stages}
    out["bitwise_equal"] = bool(np.array_equal(finals[1.0], finals[0.0]) and np.array_equal(finals[1.0], finals[0.5]))
    return out


def numpy_equal(a, b):
    import numpy as np
    return bool(np.array_equal(a, b))


SERVE_POLICY = {"policy": "oldest_first", "capacity": 1 << 19}


def served_workload(n, rate, seed=1):
    """Poisson arrivals of Euler vortex solves (sides 128..2048, 100..2000
    steps, 30% FP32 state): the continuous-batching engine against serving
    them one at a time, in arrival order, with the plain standard kernel.

    The engine's policy for the stream is "oldest_first" (rounds take the
    running problems in arrival order while they fill at most 2^19 cells):
    sharing rounds among large solves delays every one of them (processor
    sharing against first-come-first-served), so the default "all" policy,
    reported beside it, has the better saturated throughput and the worse
    tail latency."""
    from pysweep_amd import baseline, batch, hip, workloads
    stream = workloads.served_stream(n=n, rate=rate, seed=seed)
    for _, p in stream:
        p.method = "auto"  # the faster kernel per plugin class
    cells = sum(p.ic.shape[1] * p.ic.shape[2] * int(round((p.globals[1] - p.globals[0]) / p.globals[2]))
                for _, p in stream)
    out = {"requests": n, "rate_per_s": rate, "fp32_share": sum(p.precision == "fp32" for _, p in stream) / n,
           "cell_steps": cells, "engine_policy": SERVE_POLICY}

    def serve(**kw):
        batch.release_rings()  # the engine's peak HBM is its own
        eng = batch.Engine(hbm_budget=96 << 30, **kw)
        hip.synchronize()
        eng.t0 = eng.clock()
        tickets = [eng.submit(p, arrival=t) for t, p in stream]
        eng.run()
        hip.synchronize()
        lat = [t.latency for t in tickets]
        span = max(t.finished for t in tickets) - min(t.arrival for t in tickets)
        return tickets, {"p50_s": batch.percentile(lat, 50), "p99_s": batch.percentile(lat, 99),
                         "mean_s": sum(lat) / n, "solves_per_s": n / span, "gcups": cells / span / 1e9,
                         "peak_hbm_gb": eng.peak_in_use / 1e9}

    def saturated(**kw):
        eng = batch.Engine(hbm_budget=96 << 30, **kw)
        hip.synchronize()
        t0 = time.perf_counter()
        for _, p in stream:
            eng.submit(p, arrival=0.0)
        eng.run()
        hip.synchronize()
        return time.perf_counter() - t0

    tickets, out["engine"] = serve(**SERVE_POLICY)
    _, out["engine_all_policy"] = serve()
    res = baseline.serve_one_at_a_time(stream)
    lat = [f - a for a, f, _ in res]
    span = max(f for _, f, _ in res) - min(a for a, _, _ in res)
    out["one_at_a_time_plain"] = {"p50_s": batch.percentile(lat, 50), "p99_s": batch.percentile(lat, 99),
                                  "mean_s": sum(lat) / n, "solves_per_s": n / span, "gcups": cells / span / 1e9}
    # saturated: every request queued at once (throughput, no arrival gaps)
    sat = saturated(**SERVE_POLICY)
    sat_all = saturated()
    t0 = time.perf_counter()
    baseline.serve_one_at_a_time([(0.0, p) for _, p in stream])
    plain_sat = time.perf_counter() - t0
    out["saturated"] = {"engine_solves_per_s": n / sat, "engine_all_policy_solves_per_s": n / sat_all,
                        "plain_solves_per_s": n / plain_sat, "engine_over_plain": plain_sat / sat,
                        "engine_all_policy_over_plain": plain_sat / sat_all, "engine_gcups": cells / sat / 1e9,
                        "plain_gcups": cells / plain_sat / 1e9}
    by_tag = {t.problem.tag: t.result for t in tickets}
    out["bitwise_equal_results"] = sum(numpy_equal(r, by_tag[p.tag])
                                       for (_, p), (_, _, r) in zip(sorted(stream, key=lambda x: x[0]), res))
    out["engine_over_plain_solves_per_s"] = out["engine"]["solves_per_s"] / out["one_at_a_time_plain"]["solves_per_s"]
    return out


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration(nbytes=32 << 20):
    """The library's copy kernel over ``nbytes`` each way, three times (a
    transfer of known size for scaling the HBM byte counters)."""
    from pysweep_amd import hip
    hip.init(0)
    n = nbytes // 8
    src, dst = hip.DeviceArray(8 * n), hip.DeviceArray(8 * n)
    src.fill_zero()
    hip.synchronize()
    for _ in range(3):
        hip.copy_f64(dst.ptr, src.ptr, n)
    hip.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": 8 * n, "launches": 3}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported by this build "
              "(the x-strip multi-GPU decomposition is design only)", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    np, hip, ic, g, sw, ph, st, times = setup(a.n, a.b, a.nsteps, a.generic, a.slots, a.threads)
    from pysweep_amd import device, schedule
    band = isinstance(sw, device.BandRun)
    cells = a.n * a.n
    res = {}
    if a.only in ("swept", "both", "all"):
        res["swept"] = time_solves(hip, lambda s: sw.run(times, stats=s), sw.reset, a.steps, a.warmup)
    if a.only in ("standard", "both", "all"):
        stages = 2 * a.nsteps
        res["standard"] = time_solves(hip, lambda s: st.run(1, stages + 1, stats=s), st.reset, a.steps, a.warmup)
    if a.only in ("phases", "all") and band:
        res["phases"] = time_solves(hip, lambda s: ph.run(times, stats=s), ph.reset, a.steps, a.warmup)

    out = {"metric": METRIC, "unit": "GCUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    gc = {k: cells * a.nsteps / v[0] / 1e9 for k, v in res.items()}
    head = "swept" if "swept" in res else "standard"
    out["value"] = gc[head]
    out["ms_per_step"] = res[head][0] * 1e3
    out["config"] = {"workload": f"euler-vortex-{a.n}x{a.n}-b{a.b}-{a.nsteps}steps-{head}", "nx": a.n, "ny": a.n,
                     "blocksize": a.b, "operating_points": 2, "intermediate_steps": 2, "steps_per_solve": a.nsteps,
                     "method": head, "swept_kernel": "band" if band else "phases",
                     "G": times.G, "swept_stages": times.stages, "tail_stages": times.tail,
                     "phases_lds_slots": ph.spec.slots, "phases_threads": ph.spec.threads,
                     "step": "per-cell plugin step" if a.generic else "plugin level update / band sweep",
                     "standard_tile": [st.spec.bx, st.spec.by]}
    if band:
        out["config"]["band_levels"] = sw.spec.levels
        out["config"]["band_rows"] = sw.rows
    if "standard" in gc:
        out["standard_gcups"] = gc["standard"]
        out["standard_ms_per_solve"] = res["standard"][0] * 1e3
    if "swept" in gc and "standard" in gc:
        out["swept_over_standard"] = gc["swept"] / gc["standard"]
    if "phases" in gc:
        out["phases_gcups"] = gc["phases"]
        if "standard" in gc:
            out["phases_over_standard"] = gc["phases"] / gc["standard"]

    # roofline of the dominant kernel of the headline method (events of the
    # timed region, which time that kernel only)
    stats = res[head][1].as_dict()
    kern = max(stats, key=lambda k: stats[k]["ms"])
    P = times.P if head == "swept" else None
    tiles = (a.n // a.b) ** 2
    if kern == "Standard":
        cs = cells
    elif kern == "Band":
        cs = cells * sw.spec.levels
    else:
        cs = tiles * schedule.phase_cell_stages(kern, a.b, 2, P)
    avg_s = stats[kern]["avg_ms"] / 1e3
    achieved = cs * FLOPS_PER_CELL_STAGE / avg_s / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        # counted bytes apply only to the workload and kernels they were counted on
        wl = tf.get("_workload", {})
        if (wl.get("n"), wl.get("b"), wl.get("build")) == (a.n, a.b, KERNEL_BUILD) and not a.generic:
            traffic = tf.get(kern, {}).get("bytes_per_launch")
    # the Euler step is FP64 vector arithmetic (a stencil, no contraction):
    # the roof is the FP64 vector pipe, 78.6 TFLOP/s on MI355X
    out["roofline"] = {"bound": "fp64-valu", "kernel": kern, "achieved": achieved,
                       "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                       "traffic": traffic, "avg_launch_ms": stats[kern]["avg_ms"],
                       "cell_stages_per_launch": cs, "flops_per_cell_stage": FLOPS_PER_CELL_STAGE,
                       "launches_timed": stats[kern]["launches"], "timed_one_launch_in": SAMPLE_EVERY}
    # per-kernel breakdown from the warm-up solves (every launch timed)
    for m in res:
        w, nw = (res[m][2], a.warmup) if a.warmup else (res[m][1], a.steps)
        out[f"{m}_kernel_ms_per_solve"] = {k: v["ms"] / nw for k, v in w.as_dict().items() if v["launches"]}
    out["kernel_ms_per_solve"] = out[f"{head}_kernel_ms_per_solve"]
    # parity of this run: swept == standard bitwise on the final level
    if "swept" in res and "standard" in res:
        fin_sw = sw.level(times.stages + times.tail + 1)
        fin_st = st.level(2 * a.nsteps + 1)
        out["swept_equals_standard_bitwise"] = bool(np.array_equal(fin_sw, fin_st))
        if "phases" in res:
            out["phases_equals_standard_bitwise"] = bool(np.array_equal(ph.level(times.stages + times.tail + 1),
                                                                        fin_st))
    for r in {id(x): x for x in (sw, ph, st)}.values():
        r.close()
    progress(f"headline done: swept {gc.get('swept')} GCUPS, standard {gc.get('standard')} GCUPS")
    if a.large:
        out["large_instances"] = large_instances(names=tuple(a.large_cases.split(",")), steps=a.large_steps,
                                                 reals=tuple(a.large_reals.split(",")))
        if a.precision_study:
            out["precision_sensitivity"] = precision_sensitivity()
        progress("large instances done")
    if not a.no_cpu:
        out["heat_cpu_case"] = heat_cpu_case()
        progress("heat CPU case done")
    if a.batch:
        out["batched"] = batched_workload(a.batch, legs=tuple(a.batch_legs.split(",")))
        progress("batched workload done")
    if a.serve:
        out["serving"] = served_workload(a.serve, a.rate)
        out["p50_latency_s"] = out["serving"]["engine"]["p50_s"]
        out["p99_latency_s"] = out["serving"]["engine"]["p99_s"]
        progress("served stream done")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ic, g, a.cpu_stages)
        progress("CPU baseline (vectorized oracle) done")
        out["cpu_baseline"]["product_host_path"] = product_host_path(a.n, a.b)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
