"""Interface and boundary fluxes of the DGFS system on the HIP path.

Mirrors frfs/solvers/dgfs/inters.py:17-492.  An interface side is addressed
the way the reference's views address it (backends/base/types.py View):
each face flux point maps to (row = element flux point, column = element),
and here to one offset into the flux-point matrix (variable v at +32 v).
Flux points of the two sides are paired by physical location
(BaseElements._srtd_face_fpts), and the interface points are ordered so
consecutive points touch consecutive left elements (get_opt_view_perm).
"""
import numpy as np

from dgfs_hip import fr
from dgfs_hip.solvers.dgfs.initcond import DGFSMaxwellianInitCondition


def _pair_face_fpts(ele, eL, fL, eR, fR, shift):
    rowsL = ele.basis.facefpts[fL]
    rowsR = ele.basis.facefpts[fR]
    xl = ele.plocfpts[rowsL, eL]
    # shift carries the right-hand face onto the left-hand one (periodic pairs)
    xr = ele.plocfpts[rowsR, eR] + shift
    out = []
    for i, p in enumerate(xl):
        j = int(np.argmin(np.sum((xr - p) ** 2, axis=1)))
        if np.sum((xr[j] - p) ** 2) > 1e-16 * max(1.0, np.sum(p * p)):
            raise RuntimeError("unmatched interface flux points")
        out.append((rowsL[i], rowsR[j], i))
    return out


class DGFSIntInters:
    def __init__(self, be, ele, interfaces):
        self.be = be
        self.ele = ele
        recs = []
        for (eL, fL, eR, fR, shift) in interfaces:
            for rowL, rowR, k in _pair_face_fpts(ele, eL, fL, eR, fR, np.asarray(shift)):
                recs.append((k, fL, eL, rowL, eR, rowR))
        recs.sort()
        self.ninterfpts = n = len(recs)
        loff = np.array([ele.fpt_offset(r[3], r[2]) for r in recs], dtype=np.int64)
        roff = np.array([ele.fpt_offset(r[5], r[4]) for r in recs], dtype=np.int64)
        nl = np.array([ele.norm_pnorm_fpts[r[3], r[2]] for r in recs]).reshape(n, ele.ndims).T
        magnl = np.array([ele.mag_pnorm_fpts[r[3], r[2]] for r in recs])
        self._loff = be.const_array(loff, np.int64)
        self._roff = be.const_array(roff, np.int64)
        self._nl = be.const_array(np.ascontiguousarray(nl))
        self._magnl = be.const_array(magnl)
        self._recs = recs

    def comm_flux(self, vidx):
        ele, vm = self.ele, self.ele.vm
        if self.ninterfpts == 0:
            return
        m = ele._scal_fpts
        fr.intcflux(ele.ndims, self.ninterfpts, ele.nvars, vidx, m, self._loff, m, self._roff,
                    self._nl, self._magnl, vm.d_cv(), vm.vsize(), self.be.stream_ptr)


class DGFSBCInters:
    type = None
    bctype = None

    def __init__(self, be, ele, faces, cfg, cfgsect):
        self.be = be
        self.ele = ele
        self.cfg = cfg
        self.cfgsect = cfgsect
        recs = []
        for (e, f) in faces:
            for k, row in enumerate(ele.basis.facefpts[f]):
                recs.append((k, f, e, row))
        recs.sort()
        self.ninterfpts = n = len(recs)
        self._recs = recs
        self._loff = be.const_array(np.array([ele.fpt_offset(r[3], r[2]) for r in recs], dtype=np.int64),
                                    np.int64)
        nl = np.array([ele.norm_pnorm_fpts[r[3], r[2]] for r in recs]).reshape(n, ele.ndims)
        self.nl_np = nl
        self._nl = be.const_array(np.ascontiguousarray(nl.T))
        self._magnl = be.const_array(np.array([ele.mag_pnorm_fpts[r[3], r[2]] for r in recs]))
        self.ploc = np.array([ele.plocfpts[r[3], r[2]] for r in recs])
        self._f0 = None
        self._bcvals = None
        self._rmap = None
        self._U = (0.0, 0.0, 0.0)

    def _wall_maxwellian(self):
        bc = DGFSMaxwellianInitCondition(self.be, self.cfg, self.ele.vm, self.cfgsect)
        self._f0 = self.be.const_array(bc.get_init_vals())
        u = bc.unondim()
        self._U = (float(u[0, 0]), float(u[1, 0]), float(u[2, 0]))

    def update_bc(self, t, vidx):
        pass

    def comm_flux(self, t, vidx):
        ele, vm = self.ele, self.ele.vm
        if self.ninterfpts == 0:
            return
        fr.bccflux(self.bctype, ele.ndims, self.ninterfpts, ele.nvars, vidx, ele._scal_fpts, self._loff,
                   self._nl, self._magnl, self._f0, self._bcvals, self._rmap, self._U, vm.d_cv(),
                   vm.vsize(), self.be.stream_ptr)


class DGFSWallDiffuseBCInters(DGFSBCInters):
    type = "dgfs-wall-diffuse"
    bctype = fr.BC_DIFFUSE

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()
        import torch
        self._bcvals = torch.zeros(2 * max(self.ninterfpts, 1), dtype=torch.float64, device=be.device)

    def update_bc(self, t, vidx):
        ele, vm = self.ele, self.ele.vm
        if self.ninterfpts == 0:
            return
        fr.bccupdate(self.bctype, ele.ndims, self.ninterfpts, ele.nvars, vidx, ele._scal_fpts, self._loff,
                     self._nl, self._f0, self._bcvals, self._U, vm.cw(), vm.d_cv(), vm.vsize(),
                     self.be.stream_ptr)


class DGFSWallSpecularBCInters(DGFSBCInters):
    type = "dgfs-wall-specular"
    bctype = fr.BC_SPECULAR

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        norms = self.nl_np
        if not np.allclose(norms, norms[0]):
            raise NotImplementedError("specular walls are supported on flat boundaries")
        # reflection map c -> c - 2 (c.n) n (inters.py:264-275)
        cv = ele.vm.cv()
        nrm = np.zeros(3)
        nrm[:ele.ndims] = norms[0]
        cr = cv - 2 * (nrm @ cv)[None, :] * nrm[:, None]
        rmap = np.array([int(np.argmin(np.sum((cv - cr[:, j:j + 1]) ** 2, axis=0))) for j in range(cv.shape[1])])
        if not np.all(np.sort(rmap) == np.arange(len(rmap))):
            raise AssertionError("Non-unique map")
        self._rmap = be.const_array(rmap.astype(np.int32), np.int32)
        import torch
        self._bcvals = torch.zeros(ele.vm.vsize() * max(self.ninterfpts, 1), dtype=torch.float64,
                                   device=be.device)

    def update_bc(self, t, vidx):
        ele, vm = self.ele, self.ele.vm
        if self.ninterfpts == 0:
            return
        fr.bccupdate(self.bctype, ele.ndims, self.ninterfpts, ele.nvars, vidx, ele._scal_fpts, self._loff,
                     self._nl, None, self._bcvals, self._U, vm.cw(), vm.d_cv(), vm.vsize(),
                     self.be.stream_ptr)


class DGFSInletBCInters(DGFSBCInters):
    type = "dgfs-inlet"
    bctype = fr.BC_INLET

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()


class DGFSInletNormalShockBCInters(DGFSBCInters):
    type = "dgfs-inlet-normalshock"
    bctype = fr.BC_INLET_NS

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()


class DGFSOutletBCInters(DGFSBCInters):
    type = "dgfs-outlet"
    bctype = fr.BC_OUTLET


class DGFSVanishBCInters(DGFSBCInters):
    type = "dgfs-vanish"
    bctype = fr.BC_VANISH


class DGFSWallDiffuseCylBCInters(DGFSBCInters):
    """Diffuse wall whose velocity and temperature vary along it
    (inters.py:446-492, bcs/dgfs-wall-diffuse-cyl*.mako): ux, uy, uz and T
    are expressions of x, y and [constants], evaluated at every wall flux
    point; the wall emits the unnormalised Maxwellian exp(-|c - U|^2 / T)
    scaled by the per-point flux balance."""
    type = "dgfs-wall-diffuse-cyl"
    bctype = fr.BC_DIFFUSE_CYL

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        from dgfs_hip.nputil import eval_at_points
        vm = ele.vm
        n = self.ninterfpts
        ploc = self.ploc.reshape(n, ele.ndims)
        st = np.empty((4, n))
        for d, k in enumerate(("ux", "uy", "uz")):
            st[d] = eval_at_points(cfg, cfgsect, k, ploc) / vm.u0()
        st[3] = eval_at_points(cfg, cfgsect, "T", ploc) / vm.T0()
        if np.any(st[3] <= 0):
            raise ValueError(f"{cfgsect}: wall temperature must be positive")
        self.wall_state = st
        self._f0 = be.const_array(st.ravel())
        import torch
        self._bcvals = torch.zeros(2 * max(n, 1), dtype=torch.float64, device=be.device)

    def update_bc(self, t, vidx):
        ele, vm = self.ele, self.ele.vm
        if self.ninterfpts == 0:
            return
        fr.bccupdate(self.bctype, ele.ndims, self.ninterfpts, ele.nvars, vidx, ele._scal_fpts, self._loff,
                     self._nl, self._f0, self._bcvals, self._U, vm.cw(), vm.d_cv(), vm.vsize(),
                     self.be.stream_ptr)
