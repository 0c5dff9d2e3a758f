"""The kernel-provider contract of the HIP backend, on the CPU: the walk
over providers (frfs/backends/base/backend.py:161-170), pointwise
registration by module name (frfs/backends/base/kernels.py:141-173) and the
queue items kernel(*args, **kwargs) -> (kernel, args, kwargs) run by
``<<``/``runall`` (backends/base/types.py:281-334).  No kernel is launched:
the providers here are stand-ins or raise before any launch."""
import pytest


def _backend(providers):
    from dgfs_hip.backend.base import HIPBackend
    be = object.__new__(HIPBackend)
    be._providers = list(providers)
    return be


def test_walk_skips_unsuitable_providers_in_order():
    from dgfs_hip.backend.kernels import ComputeKernel, NotSuitableError
    calls = []

    class A:
        def mul(self, *a, **kw):
            calls.append("A")
            raise NotSuitableError("not for these operands")

    class B:
        def mul(self, *a, **kw):
            calls.append("B")
            return ComputeKernel(lambda: calls.append("run"))

    class C:
        def mul(self, *a, **kw):
            calls.append("C")
            return ComputeKernel(lambda: None)

    be = _backend([A(), object(), B(), C()])
    k = be.kernel("mul", 1, 2, out=3)
    assert calls == ["A", "B"]
    k.run(None)
    assert calls[-1] == "run"
    with pytest.raises(KeyError, match="'tflux' has no providers"):
        be.kernel("tflux")


def test_pointwise_registration_by_module_name():
    from dgfs_hip.backend.kernels import HIPPointwiseKernels
    pw = HIPPointwiseKernels(backend=None)
    assert not hasattr(pw, "tflux")
    pw.register("frfs.solvers.dgfs.kernels.tflux")
    pw.register("frfs.solvers.dgfs.kernels.tflux")          # same module again: no-op
    assert callable(pw.tflux)
    with pytest.raises(RuntimeError, match="re-register"):
        pw.register("other.place.tflux")
    with pytest.raises(KeyError):
        pw.register("frfs.solvers.dgfs.kernels.mpicflux")   # no HIP kernel for it


def test_queue_runs_items_with_runtime_arguments_in_order():
    from dgfs_hip.backend.base import Queue
    from dgfs_hip.backend.kernels import ComputeKernel, ComputeMetaKernel
    seen = []
    k1 = ComputeKernel(lambda varidx=0: seen.append(("k1", varidx)))
    k2 = ComputeKernel(lambda t=0.0, varidx=0: seen.append(("k2", t, varidx)))
    meta = ComputeMetaKernel([k1, k1])
    q = Queue(backend=None)
    q << k1(varidx=32)
    q << [k2(t=0.5, varidx=64), meta(varidx=7)]
    assert len(q) == 4 - 1 and not seen
    q.run()
    assert seen == [("k1", 32), ("k2", 0.5, 64), ("k1", 7), ("k1", 7)]
    assert not q
    q % k2(t=1.0)
    assert seen[-1] == ("k2", 1.0, 0)


def test_pointwise_kernels_refuse_what_they_do_not_implement():
    """A source term, a Riemann solver other than the exact upwind flux, an
    unknown boundary or velocity rows that are not one array make the
    provider unsuitable (so another provider may serve the call)."""
    import numpy as np
    import types
    from dgfs_hip.backend import kernels as K

    class T:
        def __init__(self, p):
            self.p = p

        def data_ptr(self):
            return self.p

    be = types.SimpleNamespace(stream_ptr=0)
    rows = dict(cvx=T(1000), cvy=T(1000 + 8 * 64), cvz=T(1000 + 16 * 64))
    bad = dict(cvx=T(1000), cvy=T(5000), cvz=T(9000))
    with pytest.raises(K.NotSuitableError):
        K._negdivconf(be, dict(ndims=2, nvars=4, srcex=["u[0]*2"]), [4, 4], tdivtconf=None, rcpdjac=None)
    with pytest.raises(K.NotSuitableError):
        K._intcflux(be, dict(ndims=2, nvars=4, vsize=64, rsolver="hllc"), [3], None, None, None, None, **rows)
    with pytest.raises(K.NotSuitableError):
        K._intcflux(be, dict(ndims=2, nvars=4, vsize=64), [3], None, None, None, None, **bad)
    with pytest.raises(K.NotSuitableError):
        K._bccflux(be, dict(ndims=2, nvars=4, vsize=64, bctype="dgfs-char-riem-inv"), [3], None, None, None, **rows)
    with pytest.raises(K.NotSuitableError):
        K.HIPBlasExtKernels(be).axnpby(np.zeros(1), subdims=[0])
    # suitable arguments give a kernel without launching anything
    k = K._intcflux(be, dict(ndims=2, nvars=4, vsize=64), [0], None, None, None, None, **rows)
    assert isinstance(k, K.ComputeKernel)
