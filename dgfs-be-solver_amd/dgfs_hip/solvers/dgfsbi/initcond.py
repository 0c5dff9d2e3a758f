"""Initial and boundary distributions of the multi-species solver (mirrors
frfs/solvers/dgfsbi/initcond.py:7-212).

Species p carries the number-density distribution
    f_p = n_p (m_p / (pi T))^{3/2} exp(-m_p |c - u|^2 / T)
(initcond.py:66-73) on the shared velocity grid; a wall distribution has
n_p = 1 (the `wall=True` form, :48-55).  The support check of :75-132
recomputes n_p, u and T from the discrete distribution and raises if any
differs from its input by more than 1e-5 in dimensional units.
"""
import math
from abc import ABCMeta, abstractmethod

import numpy as np


class DGFSBiInitCondition(metaclass=ABCMeta):
    def __init__(self, backend, cfg, velocitymesh, name, **kwargs):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self._nspcs = velocitymesh.nspcs()
        self._init_vals = None
        self.load_parameters(name, **kwargs)
        self.perform_precomputation()

    @abstractmethod
    def load_parameters(self, name, **kwargs):
        pass

    @abstractmethod
    def perform_precomputation(self):
        pass

    def get_init_vals(self):
        return self._init_vals

    def apply_init_vals(self, p, scal_upts_full, ele=None):
        if not 0 <= p < self._nspcs:
            raise ValueError("species index out of range")
        scal_upts_full[:, :, :] = self._init_vals[p][None, :, None]


class DGFSBiMaxwellianInitCondition(DGFSBiInitCondition):
    model = "maxwellian"

    def load_parameters(self, name, wall=False):
        vm, cfg = self.vm, self.cfg
        u0 = vm.u0()
        self._ndenini = [1.0] * self._nspcs
        if not wall:
            for p in range(self._nspcs):
                self._ndenini[p] = cfg.getfloat(name, f"nden{p + 1}") / vm.n0()
        self._Tini = cfg.getfloat(name, "T") / vm.T0()
        self._uini = np.array([cfg.getfloat(name, "ux") / u0, cfg.getfloat(name, "uy") / u0,
                               cfg.getfloat(name, "uz") / u0]).reshape(3, 1)

    def perform_precomputation(self):
        vm = self.vm
        m, cv, cw = vm.masses(), vm.cv(), vm.cw()
        self._init_vals = []
        for p in range(self._nspcs):
            f = self._ndenini[p] * (m[p] / (math.pi * self._Tini)) ** 1.5 * \
                np.exp(-m[p] * np.sum((cv - self._uini) ** 2, axis=0) / self._Tini)
            self._init_vals.append(f)
        n0, u0, T0 = vm.n0(), vm.u0(), vm.T0()
        for p in range(self._nspcs):
            f, mcw = self._init_vals[p], m[p] * cw
            rho = f.sum() * mcw
            U = np.array([(f * cv[d]).sum() * mcw / rho for d in range(3)])
            c2 = np.sum((cv - U[:, None]) ** 2, axis=0)
            T = (f * c2).sum() * (2.0 / 3.0 * mcw * m[p]) / rho
            ok = (np.allclose(self._ndenini[p] * n0, rho * n0 / m[p], atol=1e-5)
                  and np.allclose(self._uini[:, 0] * u0, U * u0, atol=1e-5)
                  and np.allclose(self._Tini * T0, T * T0, atol=1e-5))
            if not ok:
                raise ValueError("Bulk properties not conserved! See Nv,dev")

    def unondim(self):
        return self._uini
