"""Interface and boundary fluxes of the DGFS system on the HIP path.

Mirrors frfs/solvers/dgfs/inters.py:17-492.  An interface side is a view of
the element flux-point matrix, as in the reference (backends/base/types.py
View): each interface point maps to (row = element flux point, column =
element), here to one offset into the flux-point matrix (variable v at
+32 v).  Flux points of the two sides are paired by physical location
(BaseElements._srtd_face_fpts), and the interface points are ordered so
consecutive points touch consecutive left elements (get_opt_view_perm).

Each class registers its pointwise module and keeps kernel factories in
``self.kernels`` ('comm_flux', and 'update_bc' for walls that accumulate),
built from ``backend.kernel('intcflux' | 'bccflux' | 'bccupdate',
tplargs, dims, ul=..., nl=..., magnl=..., cvx/cvy/cvz=..., bnd_f0=...,
bc_vals=...)`` with the reference's argument names.
"""
import numpy as np

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
        self._scal_lhs = be.view(ele._scal_fpts, loff)
        self._scal_rhs = be.view(ele._scal_fpts, roff)
        self._nl = be.const_array(np.ascontiguousarray(nl) if n else np.zeros((ele.ndims, 1)))
        self._magnl = be.const_array(magnl if n else np.zeros(1))
        self._recs = recs
        vm = ele.vm
        be.pointwise.register("frfs.solvers.dgfs.kernels.intcflux")
        tplargs = dict(ndims=ele.ndims, nvars=ele.nvars, rsolver="rusanov", vsize=vm.vsize())
        self.kernels = {
            "comm_flux": lambda: be.kernel("intcflux", tplargs=tplargs, dims=[self.ninterfpts],
                                           ul=self._scal_lhs, ur=self._scal_rhs, magnl=self._magnl,
                                           nl=self._nl, cvx=vm.d_cvx(), cvy=vm.d_cvy(), cvz=vm.d_cvz()),
        }


class DGFSBCInters:
    type = None
    # accumulating walls have an update kernel (bccupdate)
    updates = False

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
        self._scal_lhs = be.view(ele._scal_fpts, [ele.fpt_offset(r[3], r[2]) for r in recs])
        nl = np.array([ele.norm_pnorm_fpts[r[3], r[2]] for r in recs]).reshape(n, ele.ndims)
        self.nl_np = nl
        self._nl = be.const_array(np.ascontiguousarray(nl.T) if n else np.zeros((ele.ndims, 1)))
        self._magnl = be.const_array(np.array([ele.mag_pnorm_fpts[r[3], r[2]] for r in recs]) if n
                                     else np.zeros(1))
        self.ploc = np.array([ele.plocfpts[r[3], r[2]] for r in recs])
        self._f0 = None
        self._bcvals = None
        self._rmap = None
        self._U = (0.0, 0.0, 0.0)
        self.kernels = {}
        self._set_kernels()

    def _wall_maxwellian(self):
        bc = DGFSMaxwellianInitCondition(self.be, self.cfg, self.ele.vm, self.cfgsect)
        self._f0 = self.be.const_array(bc.get_init_vals())
        u = bc.unondim()
        self._U = (float(u[0, 0]), float(u[1, 0]), float(u[2, 0]))

    def _tplargs(self):
        vm = self.ele.vm
        c = dict(Ux=self._U[0], Uy=self._U[1], Uz=self._U[2])
        return dict(ndims=self.ele.ndims, nvars=self.ele.nvars, rsolver="rusanov", c=c, bctype=self.type,
                    vsize=vm.vsize(), cw=vm.cw())

    def _set_kernels(self):
        """inters.py:232-443: register bccflux (and bccupdate) and keep the
        kernel factories.  The factories read the boundary data (inflow or
        wall distribution, velocity, balance storage, reflection map) when the
        system instantiates them, after every subclass has set it."""
        be, vm = self.be, self.ele.vm
        be.pointwise.register("frfs.solvers.dgfs.kernels.bccflux")
        cvs = lambda: dict(cvx=vm.d_cvx(), cvy=vm.d_cvy(), cvz=vm.d_cvz())
        self.kernels["comm_flux"] = lambda: be.kernel(
            "bccflux", tplargs=self._tplargs(), dims=[self.ninterfpts], ul=self._scal_lhs, magnl=self._magnl,
            nl=self._nl, bnd_f0=self._f0, bc_vals=self._bcvals, rmap=self._rmap, **cvs())
        if self.updates:
            be.pointwise.register("frfs.solvers.dgfs.kernels.bccupdate")
            self.kernels["update_bc"] = lambda: be.kernel(
                "bccupdate", tplargs=self._tplargs(), dims=[self.ninterfpts], ul=self._scal_lhs, nl=self._nl,
                bnd_f0=self._f0, bc_vals=self._bcvals, **cvs())


class DGFSWallDiffuseBCInters(DGFSBCInters):
    type = "dgfs-wall-diffuse"
    updates = True

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()
        import torch
        self._bcvals = torch.zeros(2 * max(self.ninterfpts, 1), dtype=torch.float64, device=be.device)


class DGFSWallSpecularBCInters(DGFSBCInters):
    type = "dgfs-wall-specular"
    updates = True

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


class DGFSInletBCInters(DGFSBCInters):
    type = "dgfs-inlet"

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()


class DGFSInletNormalShockBCInters(DGFSBCInters):
    type = "dgfs-inlet-normalshock"

    def __init__(self, be, ele, faces, cfg, cfgsect):
        super().__init__(be, ele, faces, cfg, cfgsect)
        self._wall_maxwellian()


class DGFSOutletBCInters(DGFSBCInters):
    type = "dgfs-outlet"


class DGFSVanishBCInters(DGFSBCInters):
    type = "dgfs-vanish"


class DGFSWallDiffuseCylBCInters(DGFSBCInters):
    """Diffuse wall whose velocity and temperature vary along it
    (inters.py:446-492, bcs/dgfs-wall-diffuse-cyl*.mako): ux, uy, uz and T
    are expressions of x, y and [constants], evaluated at every wall flux
    point; the wall emits the unnormalised Maxwellian exp(-|c - U|^2 / T)
    scaled by the per-point flux balance."""
    type = "dgfs-wall-diffuse-cyl"
    updates = True

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
