"""Initial and boundary distributions (mirrors frfs/solvers/dgfs/initcond.py:7-140)."""
import math
from abc import ABCMeta, abstractmethod

import numpy as np


class DGFSInitCondition(metaclass=ABCMeta):
    def __init__(self, backend, cfg, velocitymesh, name):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self._init_vals = None
        self.load_parameters(name)
        self.perform_precomputation()

    @abstractmethod
    def load_parameters(self, name):
        pass

    @abstractmethod
    def perform_precomputation(self):
        pass

    def get_init_vals(self):
        return self._init_vals

    def apply_init_vals(self, scal_upts_full, ele=None):
        scal_upts_full[:, :, :] = self._init_vals[None, :, None]


class DGFSMaxwellianInitCondition(DGFSInitCondition):
    model = "maxwellian"

    def load_parameters(self, name):
        vm, cfg = self.vm, self.cfg
        u0 = vm.u0()
        self._rhoini = cfg.getfloat(name, "rho") / vm.rho0()
        self._Tini = cfg.getfloat(name, "T") / vm.T0()
        self._uini = np.array([cfg.getfloat(name, "ux") / u0, cfg.getfloat(name, "uy") / u0,
                               cfg.getfloat(name, "uz") / u0]).reshape(3, 1)

    def perform_precomputation(self):
        vm = self.vm
        cv = vm.cv()
        self._init_vals = (self._rhoini / (math.pi * self._Tini) ** 1.5) * \
            np.exp(-np.sum((cv - self._uini) ** 2, axis=0) / self._Tini)
        # support check of initcond.py:59-113: moments must be reproduced to 1e-5
        f, cw = self._init_vals, vm.cw()
        rho = f.sum() * cw
        U = np.array([(f * cv[d]).sum() * cw / rho for d in range(3)])
        c2 = np.sum((cv - U[:, None]) ** 2, axis=0)
        T = (f * c2).sum() * cw * 2.0 / 3.0 / rho
        ok = (np.allclose(self._rhoini * vm.rho0(), rho * vm.rho0(), atol=1e-5)
              and np.allclose(self._uini[:, 0] * vm.u0(), U * vm.u0(), atol=1e-5)
              and np.allclose(self._Tini * vm.T0(), T * vm.T0(), atol=1e-5))
        if not ok:
            raise ValueError("Bulk properties not conserved! Check Nv, dev")

    def unondim(self):
        return self._uini


class DGFSBKWInitCondition(DGFSInitCondition):
    model = "bkw"

    def load_parameters(self, name):
        self._t0 = self.cfg.getfloat(name, "t0")

    def perform_precomputation(self):
        K = 1.0 - math.exp(-self._t0 / 6.0)
        c2 = np.sum(self.vm.cv() ** 2, axis=0)
        self._init_vals = 1 / (2 * (2 * math.pi * K) ** 1.5) * np.exp(-c2 / (2.0 * K)) * \
            ((5 * K - 3) / K + (1 - K) / (K * K) * c2)

    def unondim(self):
        raise ValueError("Not defined for bkw")
