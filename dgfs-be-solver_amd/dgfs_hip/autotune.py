"""Launch-shape autotuner of the collision operators (single-species VHS and
two-species VSS).

For a given operator (velocity grid, precision, K = M Nrho) and batch size it
times every candidate of
  * the gain algorithm: multi-pass (any grid), register-resident (FP32:
    two pencils per thread at Nv 16 / 32 with the Nyquist planes from a
    separate kernel or inside the kernel; one pencil per thread at Nv 16 / 24)
    or the slab kernel (FP64 at Nv 16 / 24 / 32);
  * the radial-group split (work items per point: 1, 2, 4, ..., Nrho);
  * the persistent grid of the multi-pass kernels (1, 2, 4 workgroups per CU,
    or the occupancy limit),
on a sample batch, keeps the fastest and caches it in a JSON file keyed by
the problem (so a run tunes once).  Every candidate computes the same Q (up to
summation order); the tuner only moves work between workgroups.
"""
import json
import os
import time

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "dgfs_hip", "tune.json")


def _platform():
    """Device name, architecture and library version: a shape tuned on another
    GPU or by another build of the library is not reused."""
    import torch
    from dgfs_hip import _lib
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(props, "gcnArchName", "")
    return f"{props.name}|{arch}|{_lib.version()}"


def _key(op, npts, platform=""):
    bucket = 1 << max(0, int(npts - 1).bit_length())
    if getattr(op, "kind", "vhs") == "vss":
        k = (f"vss-N{op.Nv}-p{op.precision}-K{op.K}-S{op.nspcs}-iso{int(bool(op.iso))}"
             f"-proj{int(bool(op.project))}-pts{bucket}")
    else:
        k = f"vhs-N{op.Nv}-p{op.precision}-K{op.K}-proj{int(bool(op.project))}-pts{bucket}"
    return f"{platform}|{k}" if platform else k


class Autotuner:
    def __init__(self, cache_path=None, repeats=2):
        self.cache_path = cache_path or os.environ.get("DGFS_TUNE_CACHE", _DEFAULT_CACHE)
        self.repeats = repeats
        self._cache = None

    def _load(self):
        if self._cache is None:
            try:
                with open(self.cache_path) as fh:
                    self._cache = json.load(fh)
            except (OSError, ValueError):
                self._cache = {}
        return self._cache

    def _store(self):
        try:
            os.makedirs(os.path.dirname(self.cache_path), exist_ok=True)
            with open(self.cache_path, "w") as fh:
                json.dump(self._cache, fh, indent=1, sort_keys=True)
        except OSError:
            pass

    @staticmethod
    def candidates(op, cus):
        if getattr(op, "kind", "vhs") == "vss":
            return Autotuner._vss_candidates(op, cus)
        algos = [1]
        if op.precision == 32 and op.Nv in (16, 24):
            algos[:0] = [4]
        if op.precision == 32 and op.Nv in (16, 32):
            algos[:0] = [2, 3, 5]
        if op.precision == 64 and op.Nv in (16, 24, 32):
            algos[:0] = [6]
        groups = [0]
        g = 1
        while g <= op.Nrho:
            groups.append(g)
            g *= 2
        out = []
        for a in algos:
            for gr in groups:
                slots = [0] if a >= 2 else [0, cus, 2 * cus, 4 * cus]
                for sl in slots:
                    out.append({"algorithm": a, "groups": gr, "slots": sl})
        return out

    @staticmethod
    def _vss_candidates(op, cus):
        """Two-species operator: the register-resident gain (and its packed
        form for isotropic scattering) where built, the slab gain for
        isotropic scattering, the multi-pass gain with its persistent-grid
        caps; radial groups 1, 2, 4, ... for each."""
        algos = [1]
        if op.precision == 32 and op.Nv in (16, 24):
            algos[:0] = [4, 6, 7] if op.iso else [4]
        if op.precision == 64 and op.iso and op.Nv in (16, 24):
            algos[:0] = [7]
        groups = [0]
        g = 1
        while g <= op.Nrho:
            groups.append(g)
            g *= 2
        out = []
        for a in algos:
            for gr in groups:
                slots = [0] if a >= 4 else [0, cus, 2 * cus, 4 * cus]
                for sl in slots:
                    out.append({"algorithm": a, "groups": gr, "slots": sl})
        return out

    @staticmethod
    def apply(op, cfg):
        op.set_algorithm(cfg["algorithm"])
        op.set_groups(cfg["groups"])
        op.set_slots(cfg["slots"])

    def tune(self, op, f, Q, npts, stream=None, force=False, log=None):
        """Pick and apply the fastest launch shape for eval_contig(f, Q, npts);
        returns (config, {candidate: ms})."""
        import torch
        from dgfs_hip import _lib
        cache = self._load()
        key = _key(op, npts, _platform())
        if not force and key in cache:
            try:
                self.apply(op, cache[key]["best"])
                return cache[key]["best"], cache[key].get("ms", {})
            except (_lib.DGFSError, KeyError, TypeError):
                # a stale or damaged entry: forget it and tune again
                del cache[key]
        cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        timings = {}
        best, best_ms = None, float("inf")
        for cfg in self.candidates(op, cus):
            self.apply(op, cfg)
            op.eval_contig(f, Q, npts, st)  # warm-up (allocations)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(self.repeats):
                op.eval_contig(f, Q, npts, st)
            torch.cuda.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / self.repeats
            name = f"a{cfg['algorithm']}-g{cfg['groups']}-s{cfg['slots']}"
            timings[name] = ms
            if log:
                log(f"  {name}: {ms:.2f} ms")
            if ms < best_ms:
                best, best_ms = cfg, ms
        self.apply(op, best)
        cache[key] = {"best": best, "ms": timings}
        self._store()
        return best, timings
