"""Project autotuner: picks launch and layout parameters by timing candidate
settings on the device, and remembers the winner per device and problem.

The reference has no counterpart (it runs one configuration of numpy and
scipy).  The tuned quantities here are the tracer's launch shape (threads per
block, persistent blocks per CU, rays per work chunk), its kernel variant
(cell-cached or per-stage gathers) and the ray-ordering brick (entry columns
grouped in bricks of 2^shift cells), which sets how coherent the field gathers
of a wavefront are.  Candidates are timed with HIP events on the current
stream; a candidate the library rejects is skipped.
"""
import json
import os
import statistics
import tempfile

import torch

from ._lib import TTError

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "turbulence_tracing_amd", "autotune.json")


def device_key(device=None):
    p = torch.cuda.get_device_properties(device or torch.cuda.current_device())
    return f"{getattr(p, 'gcnArchName', p.name)}:{p.multi_processor_count}cu"


class Autotuner:
    def __init__(self, path=None, enabled=True):
        self.path = path or os.environ.get("TT_AUTOTUNE_CACHE", DEFAULT_CACHE)
        self.enabled = enabled
        self.cache = self._load()

    def _load(self):
        try:
            with open(self.path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}

    def _save(self):
        d = os.path.dirname(self.path)
        try:
            os.makedirs(d, exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=d, suffix=".tmp")
            with os.fdopen(fd, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home keeps the result for this process only

    @staticmethod
    def key(name, signature):
        return f"{name}|{device_key()}|{json.dumps(signature, sort_keys=True)}"

    def tune(self, name, signature, candidates, run, reps=3, warmup=1, force=False):
        """Time ``run(candidate)`` for every candidate (dicts) and return the
        fastest.  ``run`` must be repeatable (same inputs, same outputs).
        Results are cached under (name, device, signature)."""
        k = self.key(name, signature)
        if not force and k in self.cache:
            return dict(self.cache[k]["best"])
        if not candidates:
            raise ValueError("autotune: no candidates")
        stream = torch.cuda.current_stream()
        results = []
        for cand in candidates:
            try:
                for _ in range(warmup):
                    run(cand)
                ms = []
                for _ in range(reps):
                    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    a.record(stream)
                    run(cand)
                    b.record(stream)
                    b.synchronize()
                    ms.append(a.elapsed_time(b))
                results.append((statistics.median(ms), cand))
            except TTError as e:
                results.append((float("inf"), {**cand, "rejected": str(e)[:200]}))
        ok = [r for r in results if r[0] != float("inf")]
        if not ok:
            raise TTError(f"autotune {name}: every candidate was rejected")
        best_ms, best = min(ok, key=lambda r: r[0])
        self.cache[k] = {"best": best, "ms": best_ms,
                         "all": [{"ms": (m if m != float("inf") else None), **c} for m, c in results]}
        if self.enabled:
            self._save()
        return dict(best)

    def report(self, name, signature):
        return self.cache.get(self.key(name, signature))
