"""Velocity mesh of the DGFS solver (mirrors frfs/solvers/dgfs/velocitymesh.py:7-161).

The grid, cell weight, radial quadrature and half-sphere set are built
exactly as the reference builds them; the per-axis velocity arrays are kept
on the device as (1, vsize) FP64 rows like the reference's const matrices.
"""
import math

import numpy as np

from dgfs_hip.quadratures import zwgj, zwglj
from dgfs_hip.sphericaldesign import get_sphquadrule


class DGFSVelocityMesh:
    R0 = 8.3144598
    NA = 6.0221409e+23

    def __init__(self, backend, cfg):
        self.backend = backend
        self.cfg = cfg
        self._construct_velocity_mesh()
        self._load_quadrature()
        self._construct_spherical_mesh()

    @classmethod
    def from_params(cls, backend, Nv, L, Nrho, M, NvBatchSize=None, quad_rule="jacobi",
                    ssrule="womersley"):
        """Mesh on [-L, L]^3 with unit non-dimensional scales (H0 = T0 = 1,
        rho0 = molarMass0 so n0 = N_A and u0 = sqrt(2 R0))."""
        from dgfs_hip.inifile import Inifile
        cfg = Inifile()
        u0 = math.sqrt(2.0 * cls.R0 / 1.0 * 1.0)
        cfg.set("non-dim", "H0", 1.0)
        cfg.set("non-dim", "T0", 1.0)
        cfg.set("non-dim", "rho0", 1.0)
        cfg.set("non-dim", "molarMass0", 1.0)
        cfg.set("constants", "Nv", Nv)
        cfg.set("constants", "NvBatchSize", NvBatchSize or Nv ** 3)
        cfg.set("constants", "Nrho", Nrho)
        cfg.set("velocity-mesh", "cmax", 0.0)
        cfg.set("velocity-mesh", "Tmax", 1.0)
        cfg.set("velocity-mesh", "dev", L)
        cfg.set("velocity-mesh", "quad-rule", quad_rule)
        cfg.set("spherical-design-rule", "ssrule", ssrule)
        cfg.set("spherical-design-rule", "M", M)
        del u0
        return cls(backend, cfg)

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
        if self._vsize < self._NvBatchSize:
            raise ValueError("NvBatchSize must not exceed Nv^3")
        self._NvBatches = -(-self._vsize // self._NvBatchSize)
        if self._NvBatches * self._NvBatchSize != self._vsize:
            raise ValueError("NvBatchSize must divide Nv^3")
        self._Nrho = cfg.getint("constants", "Nrho", int(self._Nv / 2))

        cmax = cfg.getfloat("velocity-mesh", "cmax") / self._u0
        Tmax = cfg.getfloat("velocity-mesh", "Tmax") / self._T0
        dev = cfg.getfloat("velocity-mesh", "dev")
        self._L = cmax + dev * math.sqrt(Tmax)
        self._S = self._L * 2.0 / (3.0 + math.sqrt(2.0))
        self._R = 2 * self._S

        N, L = self._Nv, self._L
        self._cw = (2.0 * L / N) ** 3
        c0 = np.linspace(-L + L / N, L - L / N, N)
        I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        self._cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[K.ravel()]])
        self._d_cv = None

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
        self._ssrule = self.cfg.get("spherical-design-rule", "ssrule", "womersley")
        self._M = self.cfg.getint("spherical-design-rule", "M")
        srule = get_sphquadrule("symmetric", rule=self._ssrule, npts=2 * self._M)
        self._sz = srule.pts[0:self._M, :]
        self._sw = 2 * np.pi / self._M

    # accessors (velocitymesh.py:137-160)
    def Nv(self): return self._Nv
    def Nrho(self): return self._Nrho
    def NvBatches(self): return self._NvBatches
    def NvBatchSize(self): return self._NvBatchSize
    def M(self): return self._M
    def L(self): return self._L
    def vsize(self): return self._vsize
    def qz(self): return self._qz
    def qw(self): return self._qw
    def cv(self): return self._cv
    def cw(self): return self._cw
    def sz(self): return self._sz
    def sw(self): return self._sw
    def H0(self): return self._H0
    def T0(self): return self._T0
    def rho0(self): return self._rho0
    def n0(self): return self._n0
    def u0(self): return self._u0
    def molarMass0(self): return self._molarMass0

    def d_cv(self):
        """Velocity components on the device, shape (3, vsize) FP64."""
        if self._d_cv is None:
            self._d_cv = self.backend.const_array(self._cv)
        return self._d_cv

    # the three rows of d_cv(), the per-axis arrays of the reference's meshes
    def d_cvx(self): return self.d_cv()[0]
    def d_cvy(self): return self.d_cv()[1]
    def d_cvz(self): return self.d_cv()[2]
