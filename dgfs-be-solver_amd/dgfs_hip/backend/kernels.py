"""Kernel providers of the HIP backend (mirrors the provider contract of
frfs/backends/base/kernels.py:10-173 and backends/cuda/{blasext,cublas,
gimmik}.py): each provider method returns a ComputeKernel whose run(queue)
launches the HIP kernel on the backend stream."""
import numpy as np

from dgfs_hip import fr


class ComputeKernel:
    def __init__(self, fn):
        self._fn = fn

    def run(self, queue=None, *args, **kwargs):
        self._fn(*args, **kwargs)


class MetaKernel(ComputeKernel):
    def __init__(self, kernels):
        self._kernels = list(kernels)

    def run(self, queue=None, *args, **kwargs):
        for k in self._kernels:
            k.run(queue, *args, **kwargs)


class HIPBlasExtKernels:
    """axnpby, copy_to_reg, copy_from_reg (backends/cuda/blasext.py:13-195)."""

    def __init__(self, backend):
        self.backend = backend

    def axnpby(self, *arr, subdims=None):
        if any(a.traits != arr[0].traits for a in arr[1:]):
            raise ValueError("Incompatible matrix types")
        n = arr[0].nrow * arr[0].leaddim
        be = self.backend

        def fn(*consts):
            fr.axnpby(n, arr, consts, be.stream_ptr)
        return ComputeKernel(fn)

    def copy_to_reg(self, full, reg):
        """run(queue, sidx, eidx, rsidx, reidx): reg[.., rsidx..reidx, ..] =
        full[.., sidx..eidx, ..] (inclusive ranges, blasext.py:77-107)."""
        be = self.backend
        nupts, vfull, neles = full.ioshape
        vreg = reg.ioshape[1]

        def fn(sidx, eidx, rsidx, reidx):
            cnt = eidx - sidx + 1
            if cnt != reidx - rsidx + 1:
                raise ValueError("copy ranges differ in length")
            fr.copy_reg(nupts, neles, full, full.leaddim, vfull, sidx, reg, reg.leaddim, vreg, rsidx,
                        cnt, be.stream_ptr)
        return ComputeKernel(fn)

    def copy_from_reg(self, reg, full):
        be = self.backend
        nupts, vfull, neles = full.ioshape
        vreg = reg.ioshape[1]

        def fn(rsidx, reidx, sidx, eidx):
            cnt = reidx - rsidx + 1
            if cnt != eidx - sidx + 1:
                raise ValueError("copy ranges differ in length")
            fr.copy_reg(nupts, neles, reg, reg.leaddim, vreg, rsidx, full, full.leaddim, vfull, sidx,
                        cnt, be.stream_ptr)
        return ComputeKernel(fn)


class HIPMFMAKernels:
    """mul(a, b, out, alpha, beta) = alpha a b + beta out with a constant
    (backends/cuda/cublas.py:105-137, gimmik.py:18-48); a is held as a
    contiguous FP64 device array."""

    def __init__(self, backend):
        self.backend = backend

    def mul(self, a, b, out, alpha=1.0, beta=0.0):
        A = a if hasattr(a, "data_ptr") else self.backend.const_array(np.asarray(a))
        m, k = A.shape
        if b.nrow != k or out.nrow != m:
            raise ValueError(f"mul: shapes {A.shape} x ({b.nrow}, ..) -> ({out.nrow}, ..)")
        ncol = b.ncol
        be = self.backend

        def fn():
            fr.mul(A, m, k, k, b, b.leaddim, out, out.leaddim, ncol, alpha, beta, be.stream_ptr)
        return ComputeKernel(fn)
