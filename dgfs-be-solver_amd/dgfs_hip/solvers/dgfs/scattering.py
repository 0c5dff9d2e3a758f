"""Scattering-model plug-ins of the DGFS solver on the HIP path.

Mirrors frfs/solvers/dgfs/scattering.py:16-286: a model is a subclass of
DGFSScatteringModel with a ``scattering_model`` class attribute, constructed
as (backend, cfg, velocitymesh), which runs load_parameters() then
perform_precomputation(); its hot entry is fs(d_arr_in, d_arr_out, elem, upt).
The HIP model adds fs_batch(arr_in, arr_out), which evaluates every solution
point of one element type in a single library call; DGFSSystem.collide uses it.

Configuration keys read (as the reference): [scattering-model] omega, dRef,
Tref; new optional keys: precision = double | single (FFT/product precision;
the registers stay FP64) and project = true | false (FP64 projection of Q on
the five collision invariants, default true for single, false for double).
"""
import ctypes as C
import math
from abc import ABCMeta, abstractmethod

import numpy as np

from dgfs_hip import _lib


class DGFSScatteringModel(metaclass=ABCMeta):
    def __init__(self, backend, cfg, velocitymesh):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self.load_parameters()
        self.perform_precomputation()

    @abstractmethod
    def load_parameters(self):
        pass

    @abstractmethod
    def perform_precomputation(self):
        pass

    @abstractmethod
    def fs(self, d_arr_in, d_arr_out, elem, upt):
        pass


class VHSOperator:
    """Owner of a libdgfs_hip VHS handle (the cuFFT-plan + tables analogue).

    Evaluates Q on caller-owned FP64 device buffers: either the AoSoA register
    of an element type (eval_aosoa) or contiguous points (eval_contig)."""

    def __init__(self, Nv, Nrho, M, L, gamma, prefac, qz, qw, sz, sw, cw,
                 precision=64, project=False):
        self._qz = np.ascontiguousarray(qz, dtype=np.float64)
        self._qw = np.ascontiguousarray(qw, dtype=np.float64)
        self._sz = np.ascontiguousarray(sz, dtype=np.float64).reshape(M, 3)
        dp = C.POINTER(C.c_double)
        desc = _lib.dgfs_vhs_desc(
            Nv=Nv, Nrho=Nrho, M=M, precision=int(precision), project=int(bool(project)),
            L=L, gamma=gamma, prefac=prefac, sw=sw, cw=cw,
            qz=self._qz.ctypes.data_as(dp), qw=self._qw.ctypes.data_as(dp),
            sz=self._sz.ctypes.data_as(dp))
        h = C.c_void_p()
        _lib.call("dgfs_vhs_create", C.byref(desc), C.byref(h))
        self._h = h
        self.Nv, self.vsize, self.precision, self.project = Nv, Nv ** 3, precision, project
        self.Nrho, self.M, self.K = Nrho, M, Nrho * M

    @classmethod
    def from_mesh(cls, vm, gamma, prefac, precision=64, project=False):
        return cls(vm.Nv(), vm.Nrho(), vm.M(), vm.L(), gamma, prefac, vm.qz(), vm.qw(),
                   vm.sz(), vm.sw(), vm.cw(), precision=precision, project=project)

    def eval_contig(self, f, Q, npts, stream=None):
        _lib.call("dgfs_vhs_eval_contig", self._h, C.c_void_p(int(f)), C.c_void_p(int(Q)),
                  int(npts), C.c_void_p(stream or 0))

    def eval_contig_noalloc(self, f, Q, npts, stream=None):
        """eval_contig without any device allocation: the workspace is the
        library's per-device scratch array (256 MiB), shared by every call on
        that GPU, so calls that overlap on one GPU race."""
        _lib.call("dgfs_vhs_eval_contig_noalloc", self._h, C.c_void_p(int(f)), C.c_void_p(int(Q)),
                  int(npts), C.c_void_p(stream or 0))

    def eval_pair(self, f, Q, npts, other, f_other, Q_other, npts_other, stream=None):
        """This handle's batch and another grid's batch in one persistent gain launch."""
        _lib.call("dgfs_vhs_eval_pair", self._h, C.c_void_p(int(f)), C.c_void_p(int(Q)), int(npts), other._h,
                  C.c_void_p(int(f_other)), C.c_void_p(int(Q_other)), int(npts_other), C.c_void_p(stream or 0))

    def eval_aosoa(self, f, Q, nupts, ldim, vsize, neles, e0, ne, stream=None):
        _lib.call("dgfs_vhs_eval", self._h, C.c_void_p(int(f)), C.c_void_p(int(Q)),
                  int(nupts), int(ldim), int(vsize), int(neles), int(e0), int(ne),
                  C.c_void_p(stream or 0))

    def set_groups(self, ngroups):
        _lib.call("dgfs_vhs_set_groups", self._h, int(ngroups))

    def set_slots(self, nslots):
        """Cap the multi-pass persistent grid (0: occupancy limit)."""
        _lib.call("dgfs_vhs_set_slots", self._h, int(nslots))

    def get_dims(self):
        """(Nv, Nrho, M, precision, L) the handle was built for."""
        v = [C.c_int() for _ in range(4)]
        L = C.c_double()
        _lib.call("dgfs_vhs_get_dims", self._h, *[C.byref(x) for x in v], C.byref(L))
        return tuple(x.value for x in v) + (L.value,)

    def get_config(self):
        g, s, t = C.c_int(), C.c_int(), C.c_int()
        _lib.call("dgfs_vhs_get_config", self._h, C.byref(g), C.byref(s), C.byref(t))
        return {"ngroups": g.value, "nslots": s.value, "threads": t.value}

    def set_algorithm(self, algorithm):
        """0 automatic, 1 multi-pass, 2 register-resident (FP32, Nv 16/32),
        3 register-resident with in-kernel Nyquist planes, 4 register-resident
        with one pencil per thread (FP32, Nv 16/24), 5 as 2 with wave-local
        in-plane transposes, 6 slab kernel (x-slabs of the field in LDS, FP64
        or FP32, Nv 16/24/32)."""
        _lib.call("dgfs_vhs_set_algorithm", self._h, int(algorithm))

    def get_algorithm(self):
        a = C.c_int()
        _lib.call("dgfs_vhs_get_algorithm", self._h, C.byref(a))
        return a.value

    def set_timing(self, on=True):
        _lib.call("dgfs_vhs_set_timing", self._h, int(bool(on)))

    def get_timing(self):
        """{kernel: (total_ms, launches)} for prep, gain, finalize, project, nyquist."""
        ms = (C.c_double * 5)()
        n = (C.c_int * 5)()
        _lib.call("dgfs_vhs_get_timing", self._h, ms, n)
        return {k: (ms[i], n[i]) for i, k in enumerate(("prep", "gain", "finalize", "project", "nyquist"))}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("dgfs_vhs_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def vhs_prefactor(omega, dRef, Tref, H0, T0, n0, alpha=1.0):
    """scattering.py:58-62."""
    invKn = H0 * math.sqrt(2.0) * math.pi * n0 * dRef * dRef * (Tref / T0) ** (omega - 0.5)
    return invKn * alpha / (2.0 ** (2 - omega + alpha) * math.gamma(2.5 - omega) * math.pi)


class DGFSVHSGLLScatteringModel(DGFSScatteringModel):
    scattering_model = "vhs-gll"

    def load_parameters(self):
        cfg = self.cfg
        omega = cfg.getfloat("scattering-model", "omega")
        self._gamma = 2.0 * (1 - omega)
        if cfg.hasopt("scattering-model", "prefactor"):
            # direct kernel prefactor (e.g. 1/(4 pi) for the BKW test)
            self._prefactor = cfg.getfloat("scattering-model", "prefactor")
            self._Kn = float("nan")
        elif cfg.hasopt("scattering-model", "Kn"):
            # prefactor = invKn alpha / (2^(3-omega) Gamma(5/2-omega) pi)
            invKn = 1.0 / cfg.getfloat("scattering-model", "Kn")
            self._Kn = 1.0 / invKn
            self._prefactor = invKn / (2.0 ** (3.0 - omega) * math.gamma(2.5 - omega) * math.pi)
        else:
            dRef = cfg.getfloat("scattering-model", "dRef")
            Tref = cfg.getfloat("scattering-model", "Tref")
            vm = self.vm
            self._prefactor = vhs_prefactor(omega, dRef, Tref, vm.H0(), vm.T0(), vm.n0())
            invKn = vm.H0() * math.sqrt(2.0) * math.pi * vm.n0() * dRef * dRef * \
                (Tref / vm.T0()) ** (omega - 0.5)
            self._Kn = 1.0 / invKn
        prec = cfg.get("scattering-model", "precision", "double")
        if prec not in ("double", "single"):
            raise ValueError("scattering-model precision must be double or single")
        self._precision = 64 if prec == "double" else 32
        self._project = cfg.getbool("scattering-model", "project",
                                    "false" if prec == "double" else "true")

    def perform_precomputation(self):
        self.op = VHSOperator.from_mesh(self.vm, self._gamma, self._prefactor,
                                        precision=self._precision, project=self._project)
        self._one_in = None
        self._one_out = None

    @property
    def prefactor(self):
        return self._prefactor

    def fs_batch(self, d_arr_in, d_arr_out, e0=0, ne=None):
        """Q at every (elem, upt) of one element type (elements [e0, e0+ne))."""
        nrow, ldim, _ = d_arr_in.traits
        vsize, neles = d_arr_in.ioshape[1:]
        ne = neles - e0 if ne is None else ne
        self.op.eval_aosoa(d_arr_in._as_parameter_, d_arr_out._as_parameter_, nrow, ldim,
                           vsize, neles, e0, ne, self.backend.stream_ptr)

    def fs(self, d_arr_in, d_arr_out, elem, upt):
        """Q at one point, as the reference's per-point call (scattering.py:227-286)."""
        import torch
        nrow, ldim, _ = d_arr_in.traits
        vsize, neles = d_arr_in.ioshape[1:]
        if self._one_in is None:
            self._one_in = torch.empty(vsize, dtype=torch.float64, device=d_arr_in.data.device)
            self._one_out = torch.empty_like(self._one_in)
        r = torch.arange(vsize, device=d_arr_in.data.device)
        col = ((elem // 32) * vsize + r) * 32 + elem % 32
        self._one_in.copy_(d_arr_in.data[upt, col])
        self.op.eval_contig(self._one_in.data_ptr(), self._one_out.data_ptr(), 1,
                            self.backend.stream_ptr)
        d_arr_out.data[upt, col] = self._one_out
