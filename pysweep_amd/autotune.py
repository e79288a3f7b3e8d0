"""Autotuner for instance-table classes: method (swept or standard), block
size, LDS window, workgroup size and persistent blocks per CU.

``Tuner.tune`` runs a small sample batch with each candidate launch shape
and keeps the fastest (cell-steps per second); every candidate computes the
same per-cell arithmetic, so the choice never changes results (checked by
the GPU tests).  Choices persist in a JSON file next to the code-object
cache; a class without a record uses the measured defaults below.
"""
import json
import os
import time

from pysweep_amd import compiler

DEFAULT_PATH = os.path.join(compiler.CACHE, "tuning.json")

#: defaults from MI355X measurements of single-problem swept launches: the
#: Euler vortex at 2048^2 ran 9.9 GCUPS with b=16 against 8.8 with b=32 (the
#: b=32 tile's 126 KB of LDS leaves one workgroup per CU); heat 8192^2 ran
#: 183 GCUPS with b=32 and 287 with b=64, which only grids that are
#: multiples of 64 can use.
DEFAULTS = {
    ("euler", "fp64"): {"b": 16},
    ("euler", "fp32"): {"b": 16},
    ("heat", "fp64"): {"b": 32},
    ("heat", "fp32"): {"b": 32},
}

CANDIDATES = {
    "euler": [{"b": 16}, {"b": 32}, {"b": 16, "method": "standard"}],
    "heat": [{"b": 16}, {"b": 32}, {"b": 64}, {"b": 32, "method": "standard"}],
}


def _key(plugin, precision, its):
    return f"{plugin}/{precision}/its{its}"


class Tuner:
    def __init__(self, path=DEFAULT_PATH):
        self.path = path
        self.records = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.records = json.load(f)

    def record(self, plugin, precision, its):
        r = self.records.get(_key(plugin, precision, its))
        if r is None:
            r = dict(DEFAULTS.get((plugin, precision), {"b": 16}))
        return r

    def block_size(self, plugin, precision, its, nx, ny):
        b = self.record(plugin, precision, its).get("b", 16)
        for cand in (b, 32, 16, 8):
            if nx % cand == 0 and ny % cand == 0 and cand <= min(nx, ny):
                return cand
        raise ValueError(f"grid {nx}x{ny} has no supported block size")

    def method(self, plugin, precision, its):
        """"swept" or "standard" for problems submitted with method="auto"."""
        from pysweep_amd import batch
        r = self.records.get(_key(plugin, precision, its), {})
        return r.get("method", batch.DEFAULT_METHOD.get((plugin, precision), "swept"))

    def config(self, plugin, precision, its, b):
        r = self.record(plugin, precision, its)
        out = {}
        if r.get("b") == b:
            for k in ("slots", "threads"):
                if k in r:
                    out[k] = r[k]
        return out

    def tune(self, plugin, precision, its, make_problems, candidates=None, save=True):
        """Time each candidate on ``make_problems()`` (a list of Problems);
        return and store the fastest."""
        from pysweep_amd import batch, hip
        trials = []
        for cand in candidates or CANDIDATES[plugin]:
            probs = make_problems()
            for p in probs:
                p.b = cand["b"]
                p.method = cand.get("method", "swept")

            class _Fixed:
                def block_size(self_, *a):
                    return cand["b"]

                def config(self_, *a):
                    return {k: v for k, v in cand.items() if k not in ("b", "method")}

            eng = batch.Engine(tuner=_Fixed(), max_rounds_ahead=1)
            for p in probs:
                eng.submit(p, arrival=0.0)
            eng.step()  # admit + compile + first round (not timed)
            hip.synchronize()
            eng.max_rounds_ahead = 1 << 20
            t0 = time.perf_counter()
            eng.run()
            hip.synchronize()
            sec = time.perf_counter() - t0
            cells = 0
            for t in eng.finished:
                nv, nx, ny = t.problem.ic.shape
                cells += nx * ny * (t.times.stages + t.times.tail) / t.problem.its
            trials.append({"method": "swept", **cand, "seconds": sec, "gcups": cells / sec / 1e9,
                           "results": {t.problem.tag: t.result for t in eng.finished}})
        best = max(trials, key=lambda r: r["gcups"])
        self.last_trials = trials
        rec = {k: v for k, v in best.items() if k in ("b", "slots", "threads", "method")}
        trials = [{k: v for k, v in t.items() if k != "results"} for t in trials]
        rec["trials"] = trials
        self.records[_key(plugin, precision, its)] = rec
        if save and self.path:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.records, f, indent=1)
        return rec
