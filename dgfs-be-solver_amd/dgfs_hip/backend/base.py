"""HIP backend: allocation, matrix types and the kernel-provider walk.

Mirrors the contract of frfs/backends/base/backend.py:28-176: kernels are
looked up by name across the providers in order, a provider that cannot
serve the arguments raises NotSuitableError and the next one is tried
(backend.py:161-170).  The providers (backend/kernels.py) are registered
when the backend is made: BLAS extensions, the MFMA operator product and
the pointwise kernels.  Queues take the reference's queue items
``kernel(*args, **kwargs) -> (kernel, args, kwargs)`` with ``<<``, ``%`` and
``runall`` (backends/base/types.py:281-334); execution is stream-ordered on
the backend stream, so a queue run is a sequence of HIP launches.
"""
import numpy as np

from dgfs_hip.backend.kernels import (HIPBlasExtKernels, HIPMFMAKernels, HIPPointwiseKernels,
                                      NotSuitableError)
from dgfs_hip.backend.types import ConstMatrix, Matrix, MatrixBank, View

__all__ = ["HIPBackend", "NotSuitableError", "Queue"]


class Queue:
    """Kernel queue: items run in insertion order on the backend stream."""

    def __init__(self, backend):
        self.backend = backend
        self._items = []

    @staticmethod
    def _as_items(items):
        if isinstance(items, tuple) and len(items) == 3 and hasattr(items[0], "run"):
            return [items]
        if hasattr(items, "run"):
            return [(items, (), {})]
        out = []
        for it in items:
            out.extend(Queue._as_items(it))
        return out

    def __lshift__(self, items):
        self._items.extend(self._as_items(items))
        return self

    def __mod__(self, items):
        self.run()
        self << items
        self.run()
        return self

    def __bool__(self):
        return bool(self._items)

    def __len__(self):
        return len(self._items)

    def run(self):
        items, self._items = self._items, []
        for kern, args, kwargs in items:
            kern.run(self, *args, **kwargs)


class HIPBackend:
    name = "hip"

    def __init__(self, cfg=None, device=None, own_stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("HIPBackend needs a visible HIP device")
        devid = 0
        if cfg is not None and cfg.hasopt("backend-hip", "device-id"):
            d = cfg.get("backend-hip", "device-id")
            devid = int(d) if d.isdigit() else 0
        if device is not None:
            devid = int(device)
        self.device = torch.device("cuda", devid)
        torch.cuda.set_device(self.device)
        self.alignb = 128
        self.soasz = 32
        prec = cfg.get("backend", "precision", "double") if cfg is not None else "double"
        if prec not in ("double",):
            raise ValueError("the distribution registers are FP64 (precision = double)")
        self.fpdtype = np.float64
        self.pointwise = HIPPointwiseKernels(self)
        self._providers = [HIPBlasExtKernels(self), HIPMFMAKernels(self), self.pointwise]
        # kernels run on torch's current stream, or on a stream of their own
        # ([backend-hip] stream = own), which HIP graph capture of the time
        # steps needs (the legacy default stream cannot be captured)
        if own_stream is None:
            own_stream = cfg is not None and cfg.get("backend-hip", "stream", "default") == "own"
        self._stream = torch.cuda.Stream(self.device) if own_stream else torch.cuda.current_stream(self.device)
        self.own_stream = bool(own_stream)

    @property
    def stream_ptr(self):
        return self._stream.cuda_stream

    def _alloc(self, shape, dtype):
        import torch
        tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
               np.dtype(np.int32): torch.int32}[np.dtype(dtype)]
        t = torch.zeros(shape, dtype=tdt, device=self.device)
        if self.own_stream:
            # the fill ran on torch's stream; the solver's stream must see it
            torch.cuda.synchronize(self.device)
        return t

    def matrix(self, ioshape, initval=None, extent=None, aliases=None, tags=set()):
        return Matrix(self, ioshape, initval, extent, aliases, tags)

    def const_matrix(self, initval, extent=None, tags=set()):
        return ConstMatrix(self, initval, extent, tags)

    def matrix_bank(self, mats, initbank=0):
        return MatrixBank(self, mats, initbank)

    def view(self, base, offsets):
        """Flux points of one interface side: base matrix + element offsets."""
        return View(self, base, offsets)

    def const_array(self, ary, dtype=np.float64):
        import torch
        return torch.as_tensor(np.ascontiguousarray(ary, dtype=dtype), device=self.device)

    def queue(self):
        return Queue(self)

    def runall(self, queues):
        for q in queues:
            q.run()

    def kernel(self, name, *args, **kwargs):
        for prov in self._providers:
            kern = getattr(prov, name, None)
            if kern is None:
                continue
            try:
                return kern(*args, **kwargs)
            except NotSuitableError:
                continue
        raise KeyError(f"'{name}' has no providers")

    def register_provider(self, prov, first=False):
        """Add a kernel provider (first=True: ahead of the built-in ones)."""
        if first:
            self._providers.insert(0, prov)
        else:
            self._providers.append(prov)

    def synchronize(self):
        import torch
        torch.cuda.synchronize(self.device)
