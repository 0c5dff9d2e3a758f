"""Kernel providers of the HIP backend.

The contract is the reference's (frfs/backends/base/kernels.py:10-173,
frfs/backends/base/backend.py:161-170):

* ``backend.kernel(name, *args, **kwargs)`` asks each provider in turn for
  a method ``name``; a provider that cannot serve the arguments raises
  NotSuitableError and the next one is tried;
* a provider method returns a kernel object; ``kernel(*args, **kwargs)``
  returns the queue item ``(kernel, args, kwargs)`` and the queue later
  calls ``kernel.run(queue, *args, **kwargs)``, so run-time scalars such as
  ``varidx`` and ``t`` arrive at run time;
* pointwise kernels are first registered by module name,
  ``backend.pointwise.register('frfs.solvers.dgfs.kernels.tflux')``, which
  adds the method ``tflux(tplargs, dims, **named_args)``.

Here every kernel is a launch of a hand-written HIP kernel of libdgfs_hip on
the backend stream.  Matrix arguments may be matrix banks: their active
register is read when the kernel runs, as the reference's banks are.

Providers and their kernels:
  HIPBlasExtKernels   axnpby, copy_to_reg, copy_from_reg   (backends/cuda/blasext.py)
  HIPMFMAKernels      mul                                  (backends/cuda/cublas.py, gimmik.py)
  HIPPointwiseKernels tflux, negdivconf, intcflux, bccflux, bccupdate
                      (frfs/solvers/dgfs/kernels/*.mako)
"""
import numpy as np

from dgfs_hip import fr


class NotSuitableError(Exception):
    pass


class ComputeKernel:
    ktype = "compute"

    def __init__(self, launch):
        self._launch = launch

    def __call__(self, *args, **kwargs):
        return self, args, kwargs

    def run(self, queue=None, *args, **kwargs):
        self._launch(*args, **kwargs)


class ComputeMetaKernel(ComputeKernel):
    """Several kernels run in order with the same run-time arguments."""

    def __init__(self, kernels):
        self._kernels = list(kernels)
        super().__init__(self._run_all)

    def _run_all(self, *args, **kwargs):
        for k in self._kernels:
            k.run(None, *args, **kwargs)


def _ptr(x):
    """Device address of a matrix, bank (its active register), view base,
    torch tensor or None, read when a kernel runs."""
    return None if x is None else fr.ptr(x)


class HIPBlasExtKernels:
    """axnpby, copy_to_reg, copy_from_reg (backends/cuda/blasext.py:13-195)."""

    def __init__(self, backend):
        self.backend = backend

    def axnpby(self, *arr, subdims=None):
        """run(queue, a0, a1, ...): x0 = a0 x0 + a1 x1 + ... over whole rows."""
        if subdims is not None:
            raise NotSuitableError("axnpby over a subset of the variables")
        if any(a.traits != arr[0].traits for a in arr[1:]):
            raise ValueError("Incompatible matrix types")
        n = arr[0].nrow * arr[0].leaddim
        be = self.backend

        def launch(*consts):
            if len(consts) != len(arr):
                raise ValueError(f"axnpby of {len(arr)} arrays needs {len(arr)} coefficients")
            fr.axnpby(n, arr, consts, be.stream_ptr)
        return ComputeKernel(launch)

    def copy_to_reg(self, full, reg):
        """run(queue, sidx, eidx, rsidx, reidx): reg[.., rsidx..reidx, ..] =
        full[.., sidx..eidx, ..] (inclusive ranges, blasext.py:77-107)."""
        be = self.backend
        nupts, vfull, neles = full.ioshape
        vreg = reg.ioshape[1]

        def launch(sidx, eidx, rsidx, reidx):
            cnt = eidx - sidx + 1
            if cnt != reidx - rsidx + 1:
                raise ValueError("copy ranges differ in length")
            fr.copy_reg(nupts, neles, full, full.leaddim, vfull, sidx, reg, reg.leaddim, vreg, rsidx,
                        cnt, be.stream_ptr)
        return ComputeKernel(launch)

    def copy_from_reg(self, reg, full):
        """run(queue, rsidx, reidx, sidx, eidx): the reverse copy (blasext.py:110-138)."""
        be = self.backend
        nupts, vfull, neles = full.ioshape
        vreg = reg.ioshape[1]

        def launch(rsidx, reidx, sidx, eidx):
            cnt = reidx - rsidx + 1
            if cnt != eidx - sidx + 1:
                raise ValueError("copy ranges differ in length")
            fr.copy_reg(nupts, neles, reg, reg.leaddim, vreg, rsidx, full, full.leaddim, vfull, sidx,
                        cnt, be.stream_ptr)
        return ComputeKernel(launch)


class HIPMFMAKernels:
    """mul(a, b, out, alpha=1, beta=0): out = alpha a b + beta out with a
    constant operator a (backends/cuda/cublas.py:105-137, gimmik.py:18-48),
    on the FP64 MFMA GEMM frfs_mul."""

    def __init__(self, backend):
        self.backend = backend

    def mul(self, a, b, out, alpha=1.0, beta=0.0):
        be = self.backend
        if isinstance(a, np.ndarray):
            a = be.const_matrix(a)
        A = a.data if hasattr(a, "ioshape") else a
        m, k = A.shape[0], (a.ncol if hasattr(a, "ncol") else A.shape[1])
        lda = a.leaddim if hasattr(a, "leaddim") else A.shape[1]
        if b.nrow != k or out.nrow != m:
            raise ValueError(f"mul: shapes ({m}, {k}) x ({b.nrow}, ..) -> ({out.nrow}, ..)")
        if b.ncol != out.ncol:
            raise ValueError("mul: operand and output differ in columns")
        ncol = b.ncol

        def launch():
            fr.mul(A, m, k, lda, b, b.leaddim, out, out.leaddim, ncol, alpha, beta, be.stream_ptr)
        return ComputeKernel(launch)


# ---------------------------------------------------------------------------
# pointwise kernels
# ---------------------------------------------------------------------------
def _cv(backend, cvx, cvy, cvz, vsize):
    """The HIP kernels read the three velocity rows from one (3, vsize) array:
    the rows the velocity mesh hands out (d_cvx, d_cvy, d_cvz) are views of it."""
    p = _ptr(cvx)
    if _ptr(cvy) != p + 8 * vsize or _ptr(cvz) != p + 16 * vsize:
        raise NotSuitableError("cvx, cvy, cvz must be the rows of one (3, vsize) device array")
    return cvx


def _tflux(backend, tplargs, dims, u, smats, f, cvx, cvy, cvz):
    """tflux.mako: f[i][v] = sum_k smats[i][k] c_k(v) u[v] at every (upt, elem);
    run(queue, varidx)."""
    ndims, nvars, vsize = tplargs["ndims"], tplargs["nvars"], tplargs["vsize"]
    nupts, neles = dims
    cv = _cv(backend, cvx, cvy, cvz, vsize)

    def launch(varidx=0):
        fr.tflux(ndims, nupts, neles, nvars, int(varidx), u, u.leaddim, smats, smats.leaddim, f, f.leaddim,
                 cv, vsize, backend.stream_ptr)
    return ComputeKernel(launch)


def _negdivconf(backend, tplargs, dims, tdivtconf, rcpdjac, ploc=None, u=None):
    """negdivconf.mako: tdivtconf = -rcpdjac tdivtconf (+ source); run(queue, t)."""
    if any(str(ex).strip() not in ("0", "0.0") for ex in tplargs.get("srcex", [])):
        raise NotSuitableError("source terms are not supported by the HIP negdivconf")
    nvars = tplargs["nvars"]
    nupts, neles = dims

    def launch(t=0.0):
        fr.negdivconf(nupts, neles, nvars, tdivtconf, tdivtconf.leaddim, rcpdjac, rcpdjac.leaddim,
                      backend.stream_ptr)
    return ComputeKernel(launch)


def _check_upwind(tplargs):
    rs = tplargs.get("rsolver", "rusanov")
    if rs != "rusanov":
        raise NotSuitableError(f"riemann-solver {rs!r}: the DGFS flux is the exact upwind (rusanov) one")


def _intcflux(backend, tplargs, dims, ul, ur, magnl, nl, cvx, cvy, cvz):
    """intcflux.mako with the exact upwind flux; ul/ur are views of the
    flux-point matrix; run(queue, varidx)."""
    _check_upwind(tplargs)
    ndims, nvars, vsize = tplargs["ndims"], tplargs["nvars"], tplargs["vsize"]
    ninter, = dims
    cv = _cv(backend, cvx, cvy, cvz, vsize)

    def launch(varidx=0):
        if ninter == 0:
            return
        fr.intcflux(ndims, ninter, nvars, int(varidx), ul.base, ul.offsets, ur.base, ur.offsets, nl, magnl, cv,
                    vsize, backend.stream_ptr)
    return ComputeKernel(launch)


# boundary types of the HIP bccflux / bccupdate kernels by the reference's names
BCTYPES = {"dgfs-wall-diffuse": fr.BC_DIFFUSE, "dgfs-wall-specular": fr.BC_SPECULAR,
           "dgfs-inlet": fr.BC_INLET, "dgfs-outlet": fr.BC_OUTLET, "dgfs-vanish": fr.BC_VANISH,
           "dgfs-inlet-normalshock": fr.BC_INLET_NS, "dgfs-wall-diffuse-cyl": fr.BC_DIFFUSE_CYL}


def _bc_args(tplargs):
    bctype = tplargs["bctype"]
    if bctype not in BCTYPES:
        raise NotSuitableError(f"boundary {bctype!r} has no HIP kernel")
    c = tplargs.get("c", {})
    U = (float(c.get("Ux", 0.0)), float(c.get("Uy", 0.0)), float(c.get("Uz", 0.0)))
    return BCTYPES[bctype], U


def _bccflux(backend, tplargs, dims, ul, magnl, nl, cvx, cvy, cvz, bnd_f0=None, bc_vals=None, ploc=None,
             rmap=None):
    """bccflux.mako + bcs/<bctype>.mako: ghost state then the upwind flux;
    bnd_f0 is the inflow/wall distribution (or the per-point wall state of the
    cylinder wall), rmap the specular reflection map; run(queue, t, varidx)."""
    _check_upwind(tplargs)
    bctype, U = _bc_args(tplargs)
    ndims, nvars, vsize = tplargs["ndims"], tplargs["nvars"], tplargs["vsize"]
    ninter, = dims
    cv = _cv(backend, cvx, cvy, cvz, vsize)

    def launch(t=0.0, varidx=0):
        if ninter == 0:
            return
        fr.bccflux(bctype, ndims, ninter, nvars, int(varidx), ul.base, ul.offsets, nl, magnl, bnd_f0, bc_vals,
                   rmap, U, cv, vsize, backend.stream_ptr)
    return ComputeKernel(launch)


def _bccupdate(backend, tplargs, dims, ul, nl, cvx, cvy, cvz, bnd_f0=None, bc_vals=None, ploc=None):
    """bccupdate.mako + bcs/<bctype>-update.mako: the diffuse walls' flux
    balance, the specular wall's stored states; run(queue, t, varidx)."""
    bctype, U = _bc_args(tplargs)
    ndims, nvars, vsize = tplargs["ndims"], tplargs["nvars"], tplargs["vsize"]
    cw = tplargs["cw"]
    ninter, = dims
    cv = _cv(backend, cvx, cvy, cvz, vsize)

    def launch(t=0.0, varidx=0):
        if ninter == 0:
            return
        fr.bccupdate(bctype, ndims, ninter, nvars, int(varidx), ul.base, ul.offsets, nl, bnd_f0, bc_vals, U, cw,
                     cv, vsize, backend.stream_ptr)
    return ComputeKernel(launch)


class HIPPointwiseKernels:
    """register(mod) then <name>(tplargs, dims, **named_args); the kernel name
    is the last component of the module path."""

    _builders = {"tflux": _tflux, "negdivconf": _negdivconf, "intcflux": _intcflux, "bccflux": _bccflux,
                 "bccupdate": _bccupdate}

    def __init__(self, backend):
        self.backend = backend
        self._mods = {}

    def register(self, mod):
        name = mod[mod.rfind(".") + 1:]
        if name in self._mods:
            if self._mods[name] != mod:
                raise RuntimeError(f'Attempt to re-register "{name}" with a different module')
            return
        if name not in self._builders:
            raise KeyError(f"no HIP kernel implements {mod!r}")
        self._mods[name] = mod
        build, be = self._builders[name], self.backend

        def kernel_meth(tplargs, dims, **kwargs):
            return build(be, tplargs, list(dims), **kwargs)
        kernel_meth.__name__ = name
        setattr(self, name, kernel_meth)
