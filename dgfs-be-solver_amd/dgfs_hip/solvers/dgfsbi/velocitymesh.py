"""Velocity mesh of the multi-species DGFS solver.

Mirrors frfs/solvers/dgfsbi/velocitymesh.py:8-160: the box half-width
L = cmax/u0 + dev sqrt(Tmax/T0), the radial support S chosen from the largest
mass ratio (:72-76), R = 2S, the radial Gauss rule on [0, R], and the
full-sphere design with sw = 4 pi / M (:129-137).  Masses are the species'
molar masses over molarMass0.
"""
import math

import numpy as np

from dgfs_hip.quadratures import zwgj, zwglj
from dgfs_hip.sphericaldesign import get_sphquadrule


class DGFSBiVelocityMesh:
    R0 = 8.3144598
    NA = 6.0221409e+23

    def __init__(self, backend, cfg, nspcs):
        self.backend = backend
        self.cfg = cfg
        self._nspcs = nspcs
        self._construct_velocity_mesh()
        self._load_quadrature()
        self._construct_spherical_mesh()

    @classmethod
    def from_params(cls, backend, Nv, L, Nrho, M, masses, quad_rule="jacobi", ssrule="womersley"):
        """Mesh on [-L, L]^3 with unit non-dimensional scales and the given
        nondimensional masses."""
        from dgfs_hip.inifile import Inifile
        cfg = Inifile()
        cfg.set("non-dim", "H0", 1.0)
        cfg.set("non-dim", "T0", 1.0)
        cfg.set("non-dim", "rho0", 1.0)
        cfg.set("non-dim", "molarMass0", 1.0)
        cfg.set("constants", "Nv", Nv)
        cfg.set("constants", "NvBatchSize", Nv ** 3)
        cfg.set("constants", "Nrho", Nrho)
        cfg.set("velocity-mesh", "cmax", 0.0)
        cfg.set("velocity-mesh", "Tmax", 1.0)
        cfg.set("velocity-mesh", "dev", L)
        cfg.set("velocity-mesh", "quad-rule", quad_rule)
        for p, m in enumerate(masses):
            cfg.set("velocity-mesh", f"molarMass{p + 1}", m)
        cfg.set("spherical-design-rule", "ssrule", ssrule)
        cfg.set("spherical-design-rule", "M", M)
        return cls(backend, cfg, len(masses))

    def _construct_velocity_mesh(self):
        cfg = self.cfg
        self._H0 = cfg.getfloat("non-dim", "H0")
        self._T0 = cfg.getfloat("non-dim", "T0")
        self._rho0 = cfg.getfloat("non-dim", "rho0")
        self._molarMass0 = cfg.getfloat("non-dim", "molarMass0")
        self._n0 = self._rho0 / self._molarMass0 * self.NA
        self._u0 = math.sqrt(2 * self.R0 / self._molarMass0 * self._T0)
        self._Nv = cfg.getint("constants", "Nv")
        self._NvBatchSize = cfg.getint("constants", "NvBatchSize", self._Nv ** 3)
        self._vsize = self._Nv ** 3
        if self._vsize % self._NvBatchSize:
            raise ValueError("NvBatchSize must divide Nv^3")
        self._NvBatches = self._vsize // self._NvBatchSize
        self._Nrho = cfg.getint("constants", "Nrho", self._Nv // 2)
        cmax = cfg.getfloat("velocity-mesh", "cmax") / self._u0
        Tmax = cfg.getfloat("velocity-mesh", "Tmax") / self._T0
        dev = cfg.getfloat("velocity-mesh", "dev")
        self._masses = [cfg.getfloat("velocity-mesh", f"molarMass{p + 1}") / self._molarMass0
                        for p in range(self._nspcs)]
        mr = max(self._masses)
        self._L = cmax + dev * math.sqrt(Tmax)
        self._S = min(self._L * 2.0 / (max(4.0 * mr / (1.0 + mr), 2.0) + math.sqrt(1.0 + mr) + 1.0),
                      self._L * 2.0 / (max(4.0 / (1.0 + mr), 2.0) + math.sqrt(1.0 + 1.0 / mr) + 1.0))
        self._R = 2 * self._S
        self._cw = (2.0 * self._L / self._Nv) ** 3
        c0 = np.linspace(-self._L + self._L / self._Nv, self._L - self._L / self._Nv, self._Nv)
        I, J, K = np.meshgrid(np.arange(self._Nv), np.arange(self._Nv), np.arange(self._Nv), indexing="ij")
        self._cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[K.ravel()]])

    def _load_quadrature(self):
        rule = self.cfg.get("velocity-mesh", "quad-rule", "jacobi")
        quads = {"jacobi": zwgj, "lobatto": zwglj}
        if rule not in quads:
            raise ValueError("Valid quads: " + str(list(quads)))
        a = self.cfg.getfloat("velocity-mesh", "quad-rule-alpha", 0.0)
        b = self.cfg.getfloat("velocity-mesh", "quad-rule-beta", 0.0)
        z, w = quads[rule](self._Nrho, a, b)
        self._qz = (self._R / 2.0) * (1.0 + z)
        self._qw = (self._R / 2.0) * w

    def _construct_spherical_mesh(self):
        self._ssrule = self.cfg.get("spherical-design-rule", "ssrule")
        self._M = self.cfg.getint("spherical-design-rule", "M")
        self._sz = get_sphquadrule("symmetric", rule=self._ssrule, npts=self._M).pts
        self._sw = 4 * math.pi / self._M

    def nspcs(self): return self._nspcs
    def masses(self): return list(self._masses)
    def Nv(self): return self._Nv
    def Nrho(self): return self._Nrho
    def NvBatches(self): return self._NvBatches
    def NvBatchSize(self): return self._NvBatchSize
    def M(self): return self._M
    def L(self): return self._L
    def vsize(self): return self._vsize
    def qz(self): return self._qz
    def qw(self): return self._qw
    def sz(self): return self._sz
    def sw(self): return self._sw
    def cv(self): return self._cv
    def cw(self): return self._cw
    def H0(self): return self._H0
    def T0(self): return self._T0
    def n0(self): return self._n0
    def u0(self): return self._u0
    def R(self): return self._R
    def S(self): return self._S
    def rho0(self): return self._rho0
    def molarMass0(self): return self._molarMass0

    def d_cv(self):
        """Velocity components on the device, shape (3, vsize) FP64 (the grid
        is shared by every species)."""
        if getattr(self, "_d_cv", None) is None:
            self._d_cv = self.backend.const_array(self._cv)
        return self._d_cv

    # the three rows of d_cv(), the per-axis arrays of the reference's meshes
    def d_cvx(self): return self.d_cv()[0]
    def d_cvy(self): return self.d_cv()[1]
    def d_cvz(self): return self.d_cv()[2]
