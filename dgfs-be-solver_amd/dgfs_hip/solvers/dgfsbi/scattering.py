"""Multi-species scattering model of the DGFS solver on the HIP path.

Mirrors frfs/solvers/dgfsbi/scattering.py:16-330: DGFSBiVSSGLLScatteringModel
(scattering_model = 'vss-gll') reads, for every ordered species pair pq,
alpha<pq>, omega<pq>, dRef<pq>, Tref<pq> and the precomputation design
(ssrulepre, Mpre), builds the b/c tables once (dgfs_vss_create) and evaluates
Q_p = sum_q Q_pq for every species at every point of an element type in one
library call (fs_batch).  fs(elem, upt, *args) keeps the reference's
signature.  Optional keys beyond the reference: prefactor<pq> (direct kernel
prefactor), precision = double | single, project = true | false.
"""
import ctypes as C
import itertools as it
import math

import numpy as np

from dgfs_hip import _lib
from dgfs_hip.sphericaldesign import get_sphquadrule

_P, _I = C.c_void_p, C.c_int
_lib.register({
    "dgfs_vss_create": [C.POINTER(_lib.dgfs_vss_desc), C.POINTER(C.c_void_p)],
    "dgfs_vss_eval": [_P, C.POINTER(_P), C.POINTER(_P), _I, _I, _I, _I, _I, _I, _P],
    "dgfs_vss_eval_contig": [_P, C.POINTER(_P), C.POINTER(_P), _I, _P],
    "dgfs_vss_destroy": [_P],
    "dgfs_vss_get_config": [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)],
    "dgfs_vss_set_groups": [_P, _I],
    "dgfs_vss_set_symmetry": [_P, _I],
    "dgfs_vss_set_slots": [_P, _I],
    "dgfs_vss_set_algorithm": [_P, _I],
    "dgfs_vss_get_algorithm": [_P, C.POINTER(_I)],
    "dgfs_vss_get_symmetry": [_P, C.POINTER(_I)],
    "dgfs_vss_set_timing": [_P, _I],
    "dgfs_vss_get_timing": [_P, C.POINTER(C.c_double), C.POINTER(_I)],
})


def vss_prefactor(mp, mq, omega, alpha, dRef, Tref, H0, T0, n0):
    """scattering.py:74-86."""
    invKn = math.sqrt(1.0 + mp / mq) * math.pi * n0 * dRef * dRef * (Tref / T0) ** (omega - 0.5) * H0
    mu = mp * mq / (mp + mq)
    return invKn * alpha / (math.sqrt(1.0 + mp / mq) * mu ** (omega - 0.5)
                            * 2.0 ** (1.0 + alpha) * math.gamma(2.5 - omega) * math.pi)


class VSSOperator:
    """Owner of a libdgfs_hip VSS handle: b/c tables of every species pair."""

    def __init__(self, Nv, Nrho, M, L, sw, cw, qz, qw, sz, szpre, masses, gamma, eta, prefac,
                 precision=64, project=False):
        S = len(masses)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (
            qz, qw, np.reshape(sz, (M, 3)), np.reshape(szpre, (-1, 3)), masses,
            np.reshape(gamma, (S, S)), np.reshape(eta, (S, S)), np.reshape(prefac, (S, S)))]
        qz_, qw_, sz_, szp_, m_, g_, e_, p_ = self._keep
        dp = C.POINTER(C.c_double)
        desc = _lib.dgfs_vss_desc(
            Nv=Nv, Nrho=Nrho, M=M, Mpre=len(szp_), nspcs=S, precision=int(precision),
            project=int(bool(project)), L=L, sw=sw, cw=cw,
            qz=qz_.ctypes.data_as(dp), qw=qw_.ctypes.data_as(dp), sz=sz_.ctypes.data_as(dp),
            szpre=szp_.ctypes.data_as(dp), masses=m_.ctypes.data_as(dp), gamma=g_.ctypes.data_as(dp),
            eta=e_.ctypes.data_as(dp), prefac=p_.ctypes.data_as(dp))
        h = C.c_void_p()
        _lib.call("dgfs_vss_create", C.byref(desc), C.byref(h))
        self._h = h
        self.nspcs, self.Nv, self.vsize = S, Nv, Nv ** 3
        # what the autotuner keys and chooses on
        self.kind, self.precision, self.project = "vss", int(precision), bool(project)
        self.Nrho, self.K = Nrho, Nrho * M
        self.iso = bool(np.all(np.asarray(eta, dtype=float) == 0.0))

    def _ptrs(self, arrs):
        if len(arrs) != self.nspcs:
            raise ValueError(f"expected {self.nspcs} species arrays, got {len(arrs)}")
        return (C.c_void_p * self.nspcs)(*[int(a) for a in arrs])

    def eval_contig(self, fs, Qs, npts, stream=None):
        _lib.call("dgfs_vss_eval_contig", self._h, self._ptrs(fs), self._ptrs(Qs), int(npts),
                  C.c_void_p(stream or 0))

    def eval_aosoa(self, fs, Qs, nupts, ldim, vsize, neles, e0, ne, stream=None):
        _lib.call("dgfs_vss_eval", self._h, self._ptrs(fs), self._ptrs(Qs), int(nupts), int(ldim),
                  int(vsize), int(neles), int(e0), int(ne), C.c_void_p(stream or 0))

    def set_groups(self, ngroups):
        _lib.call("dgfs_vss_set_groups", self._h, int(ngroups))

    def set_algorithm(self, algorithm):
        """Gain kernel: 0 automatic, 1 multi-pass, 4 register-resident (FP32,
        Nv 16 / 24), 6 packed isotropic register-resident (eta = 0 only), 7 slab
        kernel (eta = 0, Nv 16 / 24; the default there in either precision)."""
        _lib.call("dgfs_vss_set_algorithm", self._h, int(algorithm))

    def get_algorithm(self):
        """Gain kernel of the last call: 1 multi-pass, 4 register-resident,
        5 register-resident for isotropic scattering (eta = 0), 6 packed, 7 slab."""
        v = C.c_int()
        _lib.call("dgfs_vss_get_algorithm", self._h, C.byref(v))
        return v.value

    def set_slots(self, nslots):
        """Cap the persistent gain grid at nslots workgroups (0: occupancy)."""
        _lib.call("dgfs_vss_set_slots", self._h, int(nslots))

    def set_symmetry(self, on=True):
        """Antipodal sharing of the gain products (default: on when the design
        allows it); off runs the reference's per-pair direction loop."""
        _lib.call("dgfs_vss_set_symmetry", self._h, int(bool(on)))

    def get_symmetry(self):
        v = C.c_int()
        _lib.call("dgfs_vss_get_symmetry", self._h, C.byref(v))
        return bool(v.value)

    def get_config(self):
        g, s, t = C.c_int(), C.c_int(), C.c_int()
        _lib.call("dgfs_vss_get_config", self._h, C.byref(g), C.byref(s), C.byref(t))
        return {"ngroups": g.value, "nslots": s.value, "threads": t.value}

    def set_timing(self, on=True):
        _lib.call("dgfs_vss_set_timing", self._h, int(bool(on)))

    def get_timing(self):
        ms = (C.c_double * 5)()
        n = (C.c_int * 5)()
        _lib.call("dgfs_vss_get_timing", self._h, ms, n)
        return {k: (ms[i], n[i]) for i, k in enumerate(("prep", "gain", "finalize", "project"))}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("dgfs_vss_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DGFSBiScatteringModel:
    def __init__(self, backend, cfg, velocitymesh):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self._nspcs = self.vm.nspcs()
        self.load_parameters()
        self.perform_precomputation()


class DGFSBiVSSGLLScatteringModel(DGFSBiScatteringModel):
    scattering_model = "vss-gll"

    def load_parameters(self):
        cfg, vm = self.cfg, self.vm
        ssrulepre = cfg.get("scattering-model", "ssrulepre", "womersley")
        self._Mpre = cfg.getint("scattering-model", "Mpre")
        self._szpre = get_sphquadrule("symmetric", rule=ssrulepre, npts=self._Mpre).pts
        self._cases = list(it.product(range(self._nspcs), repeat=2))
        S = self._nspcs
        self._gamma = np.zeros((S, S))
        self._eta = np.zeros((S, S))
        self._prefactor = np.zeros((S, S))
        m = vm.masses()
        for p, q in self._cases:
            pq = f"{p}{q}"
            alpha = cfg.getfloat("scattering-model", "alpha" + pq)
            omega = cfg.getfloat("scattering-model", "omega" + pq)
            self._eta[p, q] = alpha - 1.0
            self._gamma[p, q] = 2.0 * (1.0 - omega)
            if cfg.hasopt("scattering-model", "prefactor" + pq):
                self._prefactor[p, q] = cfg.getfloat("scattering-model", "prefactor" + pq)
            else:
                dRef = cfg.getfloat("scattering-model", "dRef" + pq)
                Tref = cfg.getfloat("scattering-model", "Tref" + pq)
                self._prefactor[p, q] = vss_prefactor(m[p], m[q], omega, alpha, dRef, Tref,
                                                      vm.H0(), vm.T0(), vm.n0())
        prec = cfg.get("scattering-model", "precision", "double")
        if prec not in ("double", "single"):
            raise ValueError("scattering-model precision must be double or single")
        self._precision = 64 if prec == "double" else 32
        self._project = cfg.getbool("scattering-model", "project",
                                    "false" if prec == "double" else "true")

    def perform_precomputation(self):
        vm = self.vm
        self.op = VSSOperator(vm.Nv(), vm.Nrho(), vm.M(), vm.L(), vm.sw(), vm.cw(), vm.qz(), vm.qw(),
                              vm.sz(), self._szpre, vm.masses(), self._gamma, self._eta,
                              self._prefactor, precision=self._precision, project=self._project)

    @property
    def prefactor(self):
        return self._prefactor.copy()

    def fs_batch(self, arrs_in, arrs_out, e0=0, ne=None):
        """Q of every species at every (elem, upt) of one element type."""
        a0 = arrs_in[0]
        nrow, ldim, _ = a0.traits
        vsize, neles = a0.ioshape[1:]
        ne = neles - e0 if ne is None else ne
        self.op.eval_aosoa([a._as_parameter_ for a in arrs_in], [a._as_parameter_ for a in arrs_out],
                           nrow, ldim, vsize, neles, e0, ne, self.backend.stream_ptr)

    def fs(self, elem, upt, *args):
        """Reference signature (scattering.py:323-330): args = f_0..f_{S-1}, Q_0..Q_{S-1};
        evaluates every solution point of `elem` (points are independent)."""
        S = self._nspcs
        if len(args) != 2 * S:
            raise ValueError("fs needs one input and one output register per species")
        self.fs_batch(args[:S], args[S:], e0=elem, ne=1)
