"""Bindings of the FR and stage kernels of libdgfs_hip (include/dgfs_hip.h)."""
import ctypes as C

from dgfs_hip import _lib

_P, _I, _L, _D = C.c_void_p, C.c_int, C.c_long, C.c_double

_lib.register({
    "frfs_mul": [_P, _I, _I, _L, _P, _L, _P, _L, _L, _D, _D, _P],
    "dgfs_tflux": [_I, _I, _I, _I, _I, _P, _L, _P, _L, _P, _L, _P, _I, _P],
    "dgfs_negdivconf": [_I, _I, _I, _P, _L, _P, _L, _P],
    "dgfs_intcflux": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "dgfs_bccflux": [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D, _P, _I, _P],
    "dgfs_bccupdate": [_I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _D, _D, _D, _D, _P, _I, _P],
    "frfs_axnpby": [_L, _I, C.POINTER(_P), C.POINTER(_D), _P],
    "dgfs_copy_reg": [_I, _I, _P, _L, _I, _I, _P, _L, _I, _I, _I, _P],
    "dgfs_moments": [_I, _I, _I, _P, _L, _P, _D, _P, _P],
    "dgfs_residual_norms": [_I, _I, _I, _P, _P, _L, _P, _P, _P],
})

BC_DIFFUSE, BC_SPECULAR, BC_INLET, BC_OUTLET, BC_VANISH, BC_INLET_NS, BC_DIFFUSE_CYL = range(7)


def ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x._as_parameter_


def mul(A, m, k, lda, B, ldb, Cm, ldc, ncol, alpha=1.0, beta=0.0, stream=None):
    _lib.call("frfs_mul", ptr(A), m, k, lda, ptr(B), ldb, ptr(Cm), ldc, ncol, alpha, beta, stream)


def tflux(ndims, nupts, neles, nvars, vidx, u, ldu, smats, ldsm, f, ldf, cv, vsize, stream=None):
    _lib.call("dgfs_tflux", ndims, nupts, neles, nvars, vidx, ptr(u), ldu, ptr(smats), ldsm,
              ptr(f), ldf, ptr(cv), vsize, stream)


def negdivconf(nupts, neles, nvars, out, ld, rcpdjac, ldr, stream=None):
    _lib.call("dgfs_negdivconf", nupts, neles, nvars, ptr(out), ld, ptr(rcpdjac), ldr, stream)


def intcflux(ndims, nfpts, nvars, vidx, L, loff, R, roff, nl, magnl, cv, vsize, stream=None):
    _lib.call("dgfs_intcflux", ndims, nfpts, nvars, vidx, ptr(L), ptr(loff), ptr(R), ptr(roff),
              ptr(nl), ptr(magnl), ptr(cv), vsize, stream)


def bccflux(bctype, ndims, nfpts, nvars, vidx, L, loff, nl, magnl, f0, bcvals, rmap, U, cv, vsize,
            stream=None):
    _lib.call("dgfs_bccflux", bctype, ndims, nfpts, nvars, vidx, ptr(L), ptr(loff), ptr(nl),
              ptr(magnl), ptr(f0), ptr(bcvals), ptr(rmap), float(U[0]), float(U[1]), float(U[2]),
              ptr(cv), vsize, stream)


def bccupdate(bctype, ndims, nfpts, nvars, vidx, L, loff, nl, f0, bcvals, U, cw, cv, vsize, stream=None):
    _lib.call("dgfs_bccupdate", bctype, ndims, nfpts, nvars, vidx, ptr(L), ptr(loff), ptr(nl),
              ptr(f0), ptr(bcvals), float(U[0]), float(U[1]), float(U[2]), float(cw), ptr(cv),
              vsize, stream)


def axnpby(n, arrays, coeffs, stream=None):
    nv = len(arrays)
    xs = (_P * nv)(*[ptr(a) for a in arrays])
    cs = (_D * nv)(*[float(c) for c in coeffs])
    _lib.call("frfs_axnpby", n, nv, xs, cs, stream)


def copy_reg(nupts, neles, src, lds, nvsrc, s0, dst, ldd, nvdst, d0, count, stream=None):
    _lib.call("dgfs_copy_reg", nupts, neles, ptr(src), lds, nvsrc, s0, ptr(dst), ldd, nvdst, d0,
              count, stream)


def moments(nupts, neles, vsize, f, ld, cv, cw, out, stream=None):
    _lib.call("dgfs_moments", nupts, neles, vsize, ptr(f), ld, ptr(cv), cw, ptr(out), stream)


def residual_norms(nupts, neles, vsize, prev, curr, ld, out2, work, stream=None):
    _lib.call("dgfs_residual_norms", nupts, neles, vsize, ptr(prev), ptr(curr), ld, ptr(out2), ptr(work), stream)
