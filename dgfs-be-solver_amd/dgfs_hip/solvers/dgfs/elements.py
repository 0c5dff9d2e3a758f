"""FR element data and kernels of the DGFS system on the HIP path.

Mirrors frfs/solvers/dgfs/elements.py:26-164 (disu, tdisf, tdivtpcorf,
tdivtconf, negdivconf) and the geometry of
frfs/solvers/base/elements.py:189-354 (physical points, smats = |J| J^-1,
djac, physical normals at flux points).  The mapping is the linear
(straight-sided) one of the mesh nodes, evaluated directly at the target
points.

The kernels are obtained the way the reference's elements obtain them: the
pointwise modules are registered with the backend and every kernel comes
from the provider walk ``backend.kernel(name, ...)``; ``self.kernels`` maps
the reference's names to factories, and the system instantiates them once.
"""
import numpy as np


class DGFSElements:
    def __init__(self, backend, shape, nodes, vm, nvars, nregs):
        self.be = backend
        self.basis = shape
        self.vm = vm
        self.nodes = np.asarray(nodes, dtype=float)
        self.nspts, self.neles, self.ndims = self.nodes.shape
        self.nupts, self.nfpts = shape.nupts, shape.nfpts
        self.nvars = nvars
        self._sbasis = shape.shape_basis()
        self._geometry()

        be = backend
        nupts, nfpts, neles, ndims = self.nupts, self.nfpts, self.neles, self.ndims
        mk = lambda shp: be.matrix(shp, tags={"align"})
        self.scal_upts = [mk((nupts, nvars, neles)) for _ in range(nregs)]
        self.scal_upts_inb = be.matrix_bank(self.scal_upts)
        self.scal_upts_outb = be.matrix_bank(self.scal_upts)
        self._scal_fpts = mk((nfpts, nvars, neles))
        self._vect_upts = mk((ndims, nupts, nvars, neles))
        self._smats_upts = be.const_matrix(self.smat_at_np("upts"), tags={"align"})
        self._rcpdjac_upts = be.const_matrix(self.rcpdjac_at_np("upts"), tags={"align"})
        self._opmats = {}
        self._set_kernels()

    def opmat(self, expr):
        """The operator expression as a device constant matrix (cached)."""
        if expr not in self._opmats:
            self._opmats[expr] = self.be.const_matrix(self.basis.opmat(expr))
        return self._opmats[expr]

    def _set_kernels(self):
        """elements.py:26-164: disu, tdisf, tdivtpcorf, tdivtconf, negdivconf."""
        be, vm = self.be, self.vm
        be.pointwise.register("frfs.solvers.dgfs.kernels.negdivconf")
        be.pointwise.register("frfs.solvers.dgfs.kernels.tflux")
        tplargs = dict(ndims=self.ndims, nvars=self.nvars, vsize=vm.vsize())
        srctplargs = dict(ndims=self.ndims, nvars=self.nvars, srcex=[])
        self.kernels = {
            "disu": lambda: be.kernel("mul", self.opmat("M0"), self.scal_upts_inb, out=self._scal_fpts),
            "tdisf": lambda: be.kernel("tflux", tplargs=tplargs, dims=[self.nupts, self.neles],
                                       u=self.scal_upts_inb, smats=self._smats_upts, f=self._vect_upts,
                                       cvx=vm.d_cvx(), cvy=vm.d_cvy(), cvz=vm.d_cvz()),
            "tdivtpcorf": lambda: be.kernel("mul", self.opmat("M1 - M3*M2"), self._vect_upts,
                                            out=self.scal_upts_outb),
            "tdivtconf": lambda: be.kernel("mul", self.opmat("M3"), self._scal_fpts, out=self.scal_upts_outb,
                                           beta=1.0),
            "negdivconf": lambda: be.kernel("negdivconf", tplargs=srctplargs, dims=[self.nupts, self.neles],
                                            tdivtconf=self.scal_upts_outb, rcpdjac=self._rcpdjac_upts,
                                            ploc=None, u=None),
        }

    # geometry ------------------------------------------------------------
    def _map(self, pts):
        L = self._sbasis.at(pts)                                # (npts, nspts)
        dL = self._sbasis.grad_at(pts)                          # (ndims, nspts, npts)
        ploc = np.einsum("pa,aed->ped", L, self.nodes)          # (npts, neles, ndims)
        jac = np.einsum("sap,aer->pesr", dL, self.nodes)        # d x_r / d xi_s
        return ploc, jac

    def _smats_djac(self, pts):
        _, jac = self._map(pts)
        npts = len(pts)
        smats = np.empty((self.ndims, npts, self.ndims, self.neles))
        xs, xe = jac[:, :, 0, 0], jac[:, :, 1, 0]   # x_xi, x_eta
        ys, ye = jac[:, :, 0, 1], jac[:, :, 1, 1]   # y_xi, y_eta
        smats[0, :, 0], smats[0, :, 1] = ye, -xe
        smats[1, :, 0], smats[1, :, 1] = -ys, xs
        djac = xs * ye - xe * ys
        if np.any(djac < -1e-5):
            raise RuntimeError("Negative mesh Jacobians detected")
        return smats, djac

    def _pts(self, name):
        return {"upts": self.basis.upts, "fpts": self.basis.fpts}[name]

    def smat_at_np(self, name):
        return self._smats_djac(self._pts(name))[0]

    def djac_at_np(self, name):
        return self._smats_djac(self._pts(name))[1]

    def rcpdjac_at_np(self, name):
        return 1.0 / self.djac_at_np(name)

    def ploc_at_np(self, name):
        return self._map(self._pts(name))[0]

    def _geometry(self):
        smats = self.smat_at_np("fpts").transpose(1, 3, 0, 2)   # (nfpts, neles, i, k)
        pnorm = np.einsum("ijlk,il->ijk", smats, self.basis.norm_fpts)
        mag = np.sqrt(np.einsum("...i,...i", pnorm, pnorm))
        if np.any(mag < 1e-10):
            raise RuntimeError("Zero face normals detected")
        self.norm_pnorm_fpts = pnorm / mag[..., None]            # (nfpts, neles, ndims)
        self.mag_pnorm_fpts = mag                                # (nfpts, neles)
        self.plocfpts = self.ploc_at_np("fpts")                  # (nfpts, neles, ndims)

    def fpt_offset(self, fpt, e):
        """Offset of (fpt, var 0, e) in the flux-point matrix; var v adds 32 v."""
        m = self._scal_fpts
        return fpt * m.leaddim + (e // 32) * self.nvars * 32 + (e % 32)
