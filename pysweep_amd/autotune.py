"""Autotuner for instance-table classes: method (swept, standard or band), block
size b, and the occupancy of each kernel family (workgroups per CU, set
through the swept tile's LDS window and lanes, and the standard tile's
shape and lanes).

``Tuner.tune`` runs a small sample batch with each candidate and keeps the
fastest (cell-steps per second); every candidate computes the same per-cell
arithmetic, so the choice never changes results (checked by the GPU tests).
Each trial records the kernels' registers and LDS bytes and the workgroups
per CU they allow.  Choices persist in a JSON file next to the code-object
cache; ``batch.Engine`` reads them by default, and a class without a record
uses the measured defaults below.
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

#: the launch-shape keys of a candidate besides b and method
SHAPE_KEYS = ("slots", "threads", "stile", "sthreads")

#: candidates per plugin: the swept block sizes, the swept LDS window (a
#: smaller window fits more workgroups on a CU), the standard kernel with
#: tiles of 256 or 512 lanes (8 or 4 workgroups of LDS per CU for Euler), and
#: the band kernel (b only names the class there)
CANDIDATES = {
    "euler": [{"b": 16}, {"b": 32}, {"b": 16, "slots": 3}, {"b": 16, "method": "band"},
              {"b": 16, "method": "standard"},
              {"b": 16, "method": "standard", "stile": [8, 16], "sthreads": 256},
              {"b": 16, "method": "standard", "stile": [8, 32], "sthreads": 512}],
    "heat": [{"b": 16}, {"b": 32}, {"b": 64}, {"b": 32, "slots": 2}, {"b": 32, "method": "band"},
             {"b": 32, "method": "standard"},
             {"b": 32, "method": "standard", "stile": [32, 64], "sthreads": 256},
             {"b": 32, "method": "standard", "stile": [16, 128], "sthreads": 512}],
}

#: per CU on MI355X: LDS bytes, SIMDs, wavefront slots per SIMD, VGPRs per lane
CU_LDS = 160 * 1024
CU_SIMDS = 4
SIMD_WAVES = 8
SIMD_VGPRS = 512


def shape_of(cand):
    """The launch-shape keyword arguments of a candidate or record."""
    out = {k: cand[k] for k in SHAPE_KEYS if k in cand}
    if "stile" in out:
        out["stile"] = tuple(out["stile"])
    return out


def workgroups_per_cu(lds_bytes, regs, threads):
    """Resident workgroups per CU of a kernel: by LDS, and by registers
    (waves of 64 lanes share the CU's 4 SIMDs; each SIMD holds up to 8 waves
    and 512 VGPRs per lane, allocated in blocks of 8)."""
    waves = max(1, -(-threads // 64))
    by_lds = CU_LDS // max(lds_bytes, 1)
    per_simd = SIMD_WAVES if regs <= 0 else min(SIMD_WAVES, SIMD_VGPRS // (-(-regs // 8) * 8))
    by_regs = (per_simd * CU_SIMDS) // waves
    return int(min(by_lds, by_regs))


def _band_ok(plugin, precision, its):
    """Whether the class can run the band method (kernels.band_multi_capable)."""
    from pysweep_amd import batch, kernels
    from pysweep_amd.equations import path as eq_path
    py, cu64, cu32, nv, ops, _ = batch.PLUGINS[plugin]
    return kernels.band_multi_capable(eq_path(cu64 if precision == "fp64" else cu32), nv, ops, its)


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
        """"swept", "standard" or "band" for problems submitted with
        method="auto"."""
        from pysweep_amd import batch
        r = self.records.get(_key(plugin, precision, its), {})
        return r.get("method", batch.DEFAULT_METHOD.get((plugin, precision), "swept"))

    def config(self, plugin, precision, its, b):
        """Launch shapes of the class (plugin, precision, its, b): the
        record's, when it was tuned with this b."""
        r = self.record(plugin, precision, its)
        return shape_of(r) if r.get("b") == b else {}

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
                    return shape_of(cand)

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
            method = cand.get("method", "swept")
            occ = self._occupancy(eng, plugin, precision, its, cand["b"], method)
            trials.append({"method": method, **cand, "seconds": sec, "gcups": cells / sec / 1e9,
                           "workgroups_per_cu": occ,
                           "results": {t.problem.tag: t.result for t in eng.finished}})
        best = max(trials, key=lambda r: r["gcups"])
        self.last_trials = trials
        rec = {k: v for k, v in best.items() if k in ("b", "method", "workgroups_per_cu") + SHAPE_KEYS}
        trials = [{k: v for k, v in t.items() if k != "results"} for t in trials]
        rec["trials"] = trials
        self.records[_key(plugin, precision, its)] = rec
        if save and self.path:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.records, f, indent=1)
        return rec

    @staticmethod
    def _occupancy(eng, plugin, precision, its, b, method):
        """Workgroups per CU of the class's kernel that carries the work:
        the Octahedron (swept), the standard kernel or BandMulti (one
        wavefront per workgroup)."""
        from pysweep_amd import hip
        cls = eng.classes.get((plugin, precision, its, b))
        if cls is None:
            return None
        if method == "band":
            fn, threads = cls.band_multi(), cls.bspec.threads
        else:
            fn, threads = ((cls.fn["Octahedron"], cls.spec.threads) if method == "swept"
                           else (cls.fn["Standard"], cls.sspec.threads))
        info = hip.function_info(fn)
        return workgroups_per_cu(info["lds_bytes"], info["regs"], threads)

    def tune_classes(self, problems, sample=24, steps=40, seed=0, save=True):
        """Tune every (plugin, precision, its) class that occurs in
        ``problems`` on a sample of its own problems (up to ``sample`` of
        them, shortened to ``steps`` steps); returns {class key: record}."""
        import dataclasses
        import random
        groups = {}
        for p in problems:
            groups.setdefault((p.plugin, p.precision, p.its), []).append(p)
        out = {}
        rng = random.Random(seed)
        for (plugin, precision, its), ps in sorted(groups.items()):
            pick = rng.sample(ps, min(sample, len(ps)))

            def make(pick=pick):
                probs = []
                for i, p in enumerate(pick):
                    t0, dt = p.globals[0], p.globals[2]
                    g = list(p.globals)
                    g[1] = t0 + steps * dt
                    probs.append(dataclasses.replace(p, globals=g, tag=i, method="swept", b=None))
                return probs

            cands = [c for c in CANDIDATES[plugin]
                     if all(p.ic.shape[1] % c["b"] == 0 and p.ic.shape[2] % c["b"] == 0 for p in pick)
                     and (c.get("method") != "band" or _band_ok(plugin, precision, its))]
            out[_key(plugin, precision, its)] = self.tune(plugin, precision, its, make, candidates=cands, save=save)
        return out
