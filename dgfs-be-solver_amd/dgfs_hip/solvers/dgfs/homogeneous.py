"""Spatially homogeneous DGFS system: P independent points, collision only.

The reference has no homogeneous mode; its BKW relaxation is run as a
uniform state on a periodic mesh, where the FR advection vanishes to
round-off (frfs/solvers/dgfs/initcond.py:120-140 sets the state,
frfs/solvers/dgfs/system.py:140-218 advances it).  Here the points carry
no geometry and the advection term is exactly zero, so the steppers of
frfs/integrators/std/steppers.py:39-225 (dgfs-euler, dgfs-tvd-rk2) run
unchanged, stage for stage, over the batched collision operator:

* the full registers are (1, vsize, P) AoSoA matrices, the layout that
  ``fs_batch`` reads (one "element type" of P elements with one point each);
* the batch registers hold NvBatchSize velocities of every point;
* ``collide`` is one batched Q evaluation of all P points;
* ``rhs`` clears the batch output register (no advection, no boundaries);
* ``updatebc`` has nothing to update.

Configuration keys: those of the DGFS velocity mesh, scattering model and
initial condition ([soln-ics] type = bkw | maxwellian), plus
[mesh] points = P (default 1).
"""
import numpy as np

from dgfs_hip.solvers.dgfs.initcond import DGFSInitCondition
from dgfs_hip.solvers.dgfs.scattering import DGFSScatteringModel
from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
from dgfs_hip.util import subclass_where


class _PointSet:
    """The batch registers of the P points, named as the element registers
    that the steppers address (``ele.scal_upts``)."""

    def __init__(self, backend, npts, nvars, nregs):
        self.nupts, self.neles, self.nvars = 1, npts, nvars
        self.scal_upts = [backend.matrix((1, nvars, npts), tags={"align"}) for _ in range(nregs)]


class DGFSHomogeneousSystem:
    name = "dgfs-homogeneous"
    velocitymeshcls = DGFSVelocityMesh
    scatteringcls = DGFSScatteringModel

    def __init__(self, backend, cfg, nregs, npts=None, initsoln=None):
        self.backend = be = backend
        self.cfg = cfg
        self.vm = vm = self.velocitymeshcls(backend, cfg)
        self.npts = P = int(npts if npts is not None else cfg.getint("mesh", "points", 1))
        if P < 1:
            raise ValueError("a homogeneous system needs at least one point")
        self.nvars = vm.NvBatchSize()
        self.ele = _PointSet(be, P, self.nvars, nregs)
        self.ele_types = ["point"]
        smn = cfg.get("scattering-model", "type")
        self.sm = subclass_where(self.scatteringcls, scattering_model=smn)(be, cfg, vm)

        vsize = vm.vsize()
        if initsoln is not None:
            f0 = np.asarray(initsoln, dtype=float).reshape(P, vsize).T[None]
        else:
            icn = cfg.get("soln-ics", "type")
            ic = subclass_where(DGFSInitCondition, model=icn)(be, cfg, vm, "soln-ics")
            f0 = np.empty((1, vsize, P))
            ic.apply_init_vals(f0)
        self.eles_scal_upts_full = [be.matrix((1, vsize, P), f0, tags={"align"}) for _ in range(nregs)]

    # the three operations the steppers call ------------------------------
    def rhs(self, t, uinbank, foutbank, batch, varidx):
        """Advection of one velocity batch: identically zero."""
        import torch
        out = self.ele.scal_upts[foutbank]
        with torch.cuda.stream(self.backend._stream):
            out.data.zero_()

    def updatebc(self, t, uinbank, foutbank, batch, varidx):
        """No boundaries."""

    def collide(self, t, uinbank, foutbank):
        """Q at every point: one batched call."""
        f = self.eles_scal_upts_full
        self.sm.fs_batch(f[uinbank], f[foutbank])

    # host views ----------------------------------------------------------
    def points(self, bank):
        """(P, vsize) FP64 copy of full register ``bank``."""
        return np.ascontiguousarray(self.eles_scal_upts_full[bank].get()[0].T)

    def moments(self, bank):
        """(P, 15): rho, U, T, q, the six stresses and p of every point,
        from the same HIP moments kernel the spatial systems use."""
        import torch
        from dgfs_hip import fr
        f = self.eles_scal_upts_full[bank]
        nupts, vsize, neles = f.ioshape
        out = torch.empty((nupts, 15, neles), dtype=torch.float64, device=self.backend.device)
        fr.moments(nupts, neles, vsize, f, f.leaddim, self.vm.d_cv(), self.vm.cw(), out,
                   self.backend.stream_ptr)
        self.backend.synchronize()
        return out.cpu().numpy()[0].T
