"""Multi-species DGFS system on the HIP path (mirrors
frfs/solvers/dgfsbi/system.py:17-264).

Every species lives on the shared velocity grid and is advected by the same
FR operators; only the boundary distributions (per species) and the
collision operator (Q_p = sum_q Q_pq, one batched VSS call for all species)
couple them.  Registers are banked species-major as in the reference
(system.py:135-154): with nregs0 registers per species, register r of
species p is bank index p * nregs0 + r; the element batch registers follow
the same numbering.  rhs/updatebc take the species index `idx` (a string in
the reference's steppers) to pick that species' boundary kernels.
"""
from collections.abc import Mapping

import numpy as np

from dgfs_hip.shapes import get_shape
from dgfs_hip.solvers.dgfs.elements import DGFSElements
from dgfs_hip.solvers.dgfs.inters import DGFSIntInters
from dgfs_hip.solvers.dgfs.system import gen_kernels
from dgfs_hip.solvers.dgfsbi.initcond import DGFSBiInitCondition
from dgfs_hip.solvers.dgfsbi.inters import DGFSBiBCInters
from dgfs_hip.solvers.dgfsbi.scattering import DGFSBiScatteringModel
from dgfs_hip.solvers.dgfsbi.velocitymesh import DGFSBiVelocityMesh
from dgfs_hip.util import subclass_where


def _count_species(cfg):
    n = 0
    while cfg.hasopt("velocity-mesh", f"molarMass{n + 1}"):
        n += 1
    return n


class DGFSBiSystem:
    name = "dgfsbi"
    elementscls = DGFSElements
    velocitymeshcls = DGFSBiVelocityMesh
    scatteringcls = DGFSBiScatteringModel

    def __init__(self, backend, mesh, cfg, nregs, initsoln=None):
        self.backend = be = backend
        self.cfg = cfg
        self.mesh = mesh
        nspcs = _count_species(cfg)
        if nspcs < 2:
            raise ValueError("the multi-species system needs molarMass1, molarMass2, ... in [velocity-mesh]")
        if nregs % nspcs:
            raise ValueError("Should be multiple of nspcs")
        self._nspcs = nspcs
        self.nregs0 = nregs // nspcs
        self.vm = vm = self.velocitymeshcls(backend, cfg, nspcs)
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
        for bname, faces in mesh.boundaries.items():
            sect = f"soln-bcs-{bname}"
            cls = subclass_where(DGFSBiBCInters, type=cfg.get(sect, "type"))
            self.bcints.append(cls(be, self.ele, faces, cfg, sect, vm))

        djac = self.ele.djac_at_np("upts")
        self.unit_cfl_dt = float(np.sqrt(np.min(djac)) / vm.L() / self.ndims)

        smn = cfg.get("scattering-model", "type")
        self.sm = subclass_where(self.scatteringcls, scattering_model=smn)(be, cfg, vm)

        nupts, neles, vsize = self.ele.nupts, self.ele.neles, vm.vsize()
        if initsoln is not None:
            # (nupts, nspcs * vsize, neles), species-major as the reference's
            # solution files (system.py:88-135)
            if isinstance(initsoln, Mapping):
                from dgfs_hip.native import dist_fields, soln_for_system
                if initsoln.get("mesh_uuid", mesh.uuid) != mesh.uuid:
                    raise RuntimeError("Invalid solution for mesh.")
                initsoln = soln_for_system(initsoln, cfg, self.ele, dist_fields(vsize, nspcs), mesh.eletype)
            data = np.asarray(initsoln, dtype=float).reshape(nupts, nspcs * vsize, neles)
            f0 = [data[:, p * vsize:(p + 1) * vsize, :] for p in range(nspcs)]
        else:
            icn = cfg.get("soln-ics", "type")
            ic = subclass_where(DGFSBiInitCondition, model=icn)(be, cfg, vm, "soln-ics")
            f0 = []
            for p in range(nspcs):
                a = np.empty((nupts, vsize, neles))
                ic.apply_init_vals(p, a, self.ele)
                f0.append(a)
        self.eles_scal_upts_full = [be.matrix((nupts, vsize, neles), f0[p], tags={"align"})
                                    for p in range(nspcs) for _ in range(self.nregs0)]
        self.eles_scal_upts_inb_full = be.matrix_bank(self.eles_scal_upts_full)
        names = [f"{n}{p}" for p in range(nspcs) for n in ("comm_flux", "update_bc")]
        self._kernels = gen_kernels(self.ele, self.iint, self.bcints, names)
        self._queue = be.queue()

    def rhs(self, t, idx, uinbank, foutbank, batch, varidx):
        """Advection of one velocity batch of species idx (system.py:160-200)."""
        p = int(idx)
        k, q = self._kernels, self._queue
        self.ele.scal_upts_inb.active = uinbank
        self.ele.scal_upts_outb.active = foutbank
        q << k["eles", "disu"]()
        q << k["eles", "tdisf"](varidx=varidx)
        q << k["eles", "tdivtpcorf"]()
        q << k["iint", "comm_flux"](varidx=varidx)
        q << k["bcint", f"comm_flux{p}"](t=t, varidx=varidx)
        q << k["eles", "tdivtconf"]()
        q << k["eles", "negdivconf"](t=t)
        self.backend.runall([q])

    def updatebc(self, t, idx, uinbank, foutbank, batch, varidx):
        """Boundary accumulations of species idx over one batch (system.py:203-220)."""
        p = int(idx)
        k, q = self._kernels, self._queue
        self.ele.scal_upts_inb.active = uinbank
        self.ele.scal_upts_outb.active = foutbank
        q << k["eles", "disu"]()
        q << k["bcint", f"update_bc{p}"](t=t, varidx=varidx)
        self.backend.runall([q])

    def collide(self, t, regids):
        """regids = (in_0, ..., in_{S-1}, out_0, ..., out_{S-1}) full-register
        bank indices (system.py:223-235): Q_p = sum_q Q_pq at every point."""
        S = self._nspcs
        if len(regids) != 2 * S:
            raise ValueError("collide needs one input and one output register per species")
        fs = self.eles_scal_upts_full
        self.sm.fs_batch([fs[r] for r in regids[:S]], [fs[r] for r in regids[S:]])

    def mixture_moments(self, bank=0):
        """Per species the dimensional moment fields of
        plugins/dgfsmomwriterbi.py:133-400 from register `bank` of every
        species: mass density, bulk velocity, species temperature (about the
        species' own velocity), heat flux and pressure tensor about the
        mixture velocity, p and number density; 2D meshes take the mixture's
        z velocity as 0, as the reference's 2D writer does.  Returns a list
        over species of {name: (nupts, neles)}.

        The raw sums come from the HIP moments kernel (central moments about
        each species' own velocity); the shift to the mixture velocity W is
        exact algebra, with d = U_p - W and M = sum f c' c' cw:
        sum f (c-W)(c-W) cw = M + n d d,
        sum f (c-W)|c-W|^2 cw = q + 2 M d + d tr M + n d |d|^2."""
        vm = self.vm
        S, nd = self._nspcs, self.ndims
        raw = self.moments([p * self.nregs0 + bank for p in range(S)])   # (nupts, 15, neles) each
        m = vm.masses()
        rho0, u0, T0, M0 = vm.rho0(), vm.u0(), vm.T0(), vm.molarMass0()
        n = [r[:, 0] for r in raw]
        U = [r[:, 1:4] for r in raw]                                        # (nupts, 3, neles)
        rho = [m[p] * n[p] for p in range(S)]                              # nondimensional
        tot = sum(rho)
        W = sum(rho[p][:, None] * U[p] for p in range(S)) / tot[:, None]
        if nd < 3:
            W[:, 2] = 0.0
        iu = {(0, 0): 8, (0, 1): 9, (0, 2): 10, (1, 1): 11, (1, 2): 12, (2, 2): 13}
        names = "xyz"
        out = []
        for p in range(S):
            r = raw[p]
            Mc = np.empty(r.shape[:1] + (3, 3) + r.shape[2:])
            for (i, j), k in iu.items():
                Mc[:, i, j] = Mc[:, j, i] = 0.5 * r[:, k]
            # species temperature about (Ux, Uy, 0) in 2D: shifting the
            # centre by Uz adds n Uz^2 to the sum of |c'|^2
            uz = U[p][:, 2] if nd < 3 else 0.0
            T = m[p] * (r[:, 4] + (2.0 / 3.0) * uz ** 2)
            trM = Mc[:, 0, 0] + Mc[:, 1, 1] + Mc[:, 2, 2]
            # M and q are about the species' own velocity: shift them to W
            e = U[p] - W
            e2 = (e ** 2).sum(1)
            Me = np.einsum("aijb,ajb->aib", Mc, e)
            qW = r[:, 5:8] + 2 * Me + e * trM[:, None] + n[p][:, None] * e * e2[:, None]
            MW = Mc + n[p][:, None, None] * e[:, :, None] * e[:, None, :]
            f = {"rho": rho[p] * rho0}
            for i in range(nd):
                f["U" + names[i]] = U[p][:, i] * u0
            f["T"] = T * T0
            for i in range(nd):
                f["Q" + names[i]] = m[p] ** 2 * qW[:, i] * 0.5 * rho0 * u0 ** 3
            for i in range(nd):
                for j in range(i, nd):
                    f["P" + names[i] + names[j]] = 2 * m[p] ** 2 * MW[:, i, j] * 0.5 * rho0 * u0 ** 2
            # the reference's p and number density (dgfsmomwriterbi.py:248-256)
            f["p"] = (m[p] * vm.R0 / M0) * f["rho"] * f["T"]
            f["nden"] = (vm.NA / m[p] / M0) * f["rho"]
            out.append(f)
        return out

    def soln_arrays(self, bank=0):
        """{etype: (nupts, nspcs * vsize, neles)}: register `bank` of every
        species side by side (dgfsdistwriterbi.py:59-76)."""
        return {self.mesh.eletype: np.concatenate(
            [self.species_register(p, bank).get() for p in range(self._nspcs)], axis=1)}

    def dist_fields(self):
        from dgfs_hip.native import dist_fields
        return dist_fields(self.vm.vsize(), self._nspcs)

    def species_register(self, p, r):
        return self.eles_scal_upts_full[p * self.nregs0 + r]

    def moments(self, bank_per_species):
        """Per species (nupts, 15, neles) moments of the number-density
        distribution (plugins/dgfsmomwriterstd.py:14-214 per species, mass
        weights left to the caller); bank_per_species[p] is the bank index."""
        import torch
        from dgfs_hip import fr
        out = []
        for p, bank in enumerate(bank_per_species):
            f = self.eles_scal_upts_full[bank]
            nupts, vsize, neles = f.ioshape
            m = torch.empty((nupts, 15, neles), dtype=torch.float64, device=self.backend.device)
            fr.moments(nupts, neles, vsize, f, f.leaddim, self.vm.d_cv(), self.vm.cw(), m,
                       self.backend.stream_ptr)
            self.backend.synchronize()  # the kernel may run on the backend's own stream
            out.append(m.cpu().numpy())
        return out
