"""DGFS system on the HIP path (mirrors frfs/solvers/dgfs/system.py:17-231).

Holds the velocity mesh, one element type, its interior and boundary
interfaces, the scattering model and the full-velocity registers
(nupts, vsize, neles).  The element registers hold one velocity batch of
NvBatchSize velocities; rhs/updatebc act on the batch starting at varidx,
collide on the full registers, exactly as the reference's steppers expect.

Kernels are instantiated once from the factories of the elements and
interfaces (frfs/solvers/base/system.py:167-183, ``_gen_kernels``) and
run through a backend queue in the reference's order.
"""
from collections.abc import Mapping

import numpy as np

from dgfs_hip.shapes import get_shape
from dgfs_hip.solvers.dgfs.elements import DGFSElements
from dgfs_hip.solvers.dgfs.initcond import DGFSInitCondition
from dgfs_hip.solvers.dgfs.inters import DGFSBCInters, DGFSIntInters
from dgfs_hip.solvers.dgfs.scattering import DGFSScatteringModel
from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
from dgfs_hip.util import subclass_where


class KernelList(list):
    """Kernels of one name across several objects: calling the list gives
    the queue items of all of them (the reference's proxylist)."""

    def __call__(self, *args, **kwargs):
        return [k(*args, **kwargs) for k in self]


def gen_kernels(ele, iint, bcints, bcnames=("comm_flux", "update_bc")):
    """{(kind, name): kernel or KernelList}: every factory of the elements and
    interfaces instantiated once (frfs/solvers/base/system.py:167-183)."""
    out = {("eles", n): f() for n, f in ele.kernels.items()}
    out["iint", "comm_flux"] = KernelList([iint.kernels["comm_flux"]()])
    for n in bcnames:
        out["bcint", n] = KernelList(b.kernels[n]() for b in bcints if n in b.kernels)
    return out


class DGFSSystem:
    name = "dgfs"
    elementscls = DGFSElements
    velocitymeshcls = DGFSVelocityMesh
    scatteringcls = DGFSScatteringModel

    def __init__(self, backend, mesh, cfg, nregs, initsoln=None):
        self.backend = be = backend
        self.cfg = cfg
        self.mesh = mesh
        self.vm = vm = self.velocitymeshcls(backend, cfg)
        order = cfg.getint("solver", "order")
        et = mesh.eletype
        upts_rule = cfg.get(f"solver-elements-{et}", "soln-pts",
                            "gauss-legendre" if et == "quad" else "williams-shunn")
        fpts_rule = cfg.get("solver-interfaces-line", "flux-pts", "gauss-legendre")
        shape = get_shape(et, order, upts_rule, fpts_rule)
        self.nvars = vm.NvBatchSize()
        self.ele = self.elementscls(be, shape, mesh.nodes, vm, self.nvars, nregs)
        self.ele_types = [et]
        self.ndims = self.ele.ndims
        self.iint = DGFSIntInters(be, self.ele, mesh.interfaces)
        self.bcints = []
        for name, faces in mesh.boundaries.items():
            sect = f"soln-bcs-{name}"
            bctype = cfg.get(sect, "type")
            cls = subclass_where(DGFSBCInters, type=bctype)
            self.bcints.append(cls(be, self.ele, faces, cfg, sect))

        # unit-CFL time step (system.py:50-66)
        djac = self.ele.djac_at_np("upts")
        self.unit_cfl_dt = float(np.sqrt(np.min(djac)) / vm.L() / self.ndims)

        smn = cfg.get("scattering-model", "type")
        self.sm = subclass_where(self.scatteringcls, scattering_model=smn)(be, cfg, vm)

        # full-velocity registers
        nupts, neles, vsize = self.ele.nupts, self.ele.neles, vm.vsize()
        if isinstance(initsoln, Mapping):
            # a solution file (system.py:79-108): same fields and mesh,
            # interpolated onto this order's points
            from dgfs_hip.native import dist_fields, soln_for_system
            if initsoln.get("mesh_uuid", mesh.uuid) != mesh.uuid:
                raise RuntimeError("Invalid solution for mesh.")
            f0 = soln_for_system(initsoln, cfg, self.ele, dist_fields(vsize), et)
        elif initsoln is not None:
            f0 = np.asarray(initsoln, dtype=float).reshape(nupts, vsize, neles)
        else:
            icn = cfg.get("soln-ics", "type")
            ic = subclass_where(DGFSInitCondition, model=icn)(be, cfg, vm, "soln-ics")
            f0 = np.empty((nupts, vsize, neles))
            ic.apply_init_vals(f0, self.ele)
        self.eles_scal_upts_full = [be.matrix((nupts, vsize, neles), f0, tags={"align"})
                                    for _ in range(nregs)]
        self.eles_scal_upts_inb_full = be.matrix_bank(self.eles_scal_upts_full)
        self._kernels = gen_kernels(self.ele, self.iint, self.bcints)
        self._queue = be.queue()

    def rhs(self, t, uinbank, foutbank, batch, varidx):
        """Advection of one velocity batch: fout = -div F(uin) (system.py:140-177)."""
        k, q = self._kernels, self._queue
        self.ele.scal_upts_inb.active = uinbank
        self.ele.scal_upts_outb.active = foutbank
        q << k["eles", "disu"]()
        q << k["eles", "tdisf"](varidx=varidx)
        q << k["eles", "tdivtpcorf"]()
        q << k["iint", "comm_flux"](varidx=varidx)
        q << k["bcint", "comm_flux"](t=t, varidx=varidx)
        q << k["eles", "tdivtconf"]()
        q << k["eles", "negdivconf"](t=t)
        self.backend.runall([q])

    def updatebc(self, t, uinbank, foutbank, batch, varidx):
        """Boundary accumulations over one velocity batch (system.py:180-196)."""
        k, q = self._kernels, self._queue
        self.ele.scal_upts_inb.active = uinbank
        self.ele.scal_upts_outb.active = foutbank
        q << k["eles", "disu"]()
        q << k["bcint", "update_bc"](t=t, varidx=varidx)
        self.backend.runall([q])

    def collide(self, t, uinbank, foutbank):
        """Q at every solution point: one batched call (system.py:199-218)."""
        fsoln = self.eles_scal_upts_full
        self.sm.fs_batch(fsoln[uinbank], fsoln[foutbank])

    def soln_arrays(self, bank):
        """{etype: (nupts, vsize, neles)} of full register `bank`, the
        arrays of a distribution file (dgfsdistwriterstd.py:59-61)."""
        return {self.ele_types[0]: self.eles_scal_upts_full[bank].get()}

    def dist_fields(self):
        from dgfs_hip.native import dist_fields
        return dist_fields(self.vm.vsize())

    def moments(self, bank):
        """(nupts, 15, neles): rho, Ux, Uy, Uz, T, qx, qy, qz, Pxx, Pxy, Pxz,
        Pyy, Pyz, Pzz, p (plugins/dgfsmomwriterstd.py:14-214)."""
        import torch
        from dgfs_hip import fr
        f = self.eles_scal_upts_full[bank]
        nupts, vsize, neles = f.ioshape
        out = torch.empty((nupts, 15, neles), dtype=torch.float64, device=self.backend.device)
        fr.moments(nupts, neles, vsize, f, f.leaddim, self.vm.d_cv(), self.vm.cw(), out,
                   self.backend.stream_ptr)
        self.backend.synchronize()  # the kernel may run on the backend's own stream
        return out.cpu().numpy()
