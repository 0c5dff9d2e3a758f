"""FR element data and kernels of the DGFS system on the HIP path.

Mirrors frfs/solvers/dgfs/elements.py:26-164 (disu, tdisf, tdivtpcorf,
tdivtconf, negdivconf, invmass) and the geometry of
frfs/solvers/base/elements.py:189-354 (physical points, smats = |J| J^-1,
djac, physical normals at flux points).  The mapping is the linear
(straight-sided) one of the mesh nodes, evaluated directly at the target
points.
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

        self._M0 = be.const_array(shape.m0)
        self._M132 = be.const_array(shape.opmat("M1 - M3*M2"))
        self._M3 = be.const_array(shape.m3)
        self._smats_upts = be.const_matrix(self.smat_at_np("upts"), tags={"align"})
        self._rcpdjac_upts = be.const_matrix(self.rcpdjac_at_np("upts"), tags={"align"})

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

    # kernels (elements.py:56-164) -----------------------------------------
    def disu(self, uin):
        u = self.scal_upts[uin]
        self._mul(self._M0, u, self._scal_fpts)

    def tdisf(self, uin, vidx):
        from dgfs_hip import fr
        u = self.scal_upts[uin]
        fr.tflux(self.ndims, self.nupts, self.neles, self.nvars, vidx, u, u.leaddim,
                 self._smats_upts, self._smats_upts.leaddim, self._vect_upts,
                 self._vect_upts.leaddim, self.vm.d_cv(), self.vm.vsize(), self.be.stream_ptr)

    def tdivtpcorf(self, fout):
        self._mul(self._M132, self._vect_upts, self.scal_upts[fout])

    def tdivtconf(self, fout):
        self._mul(self._M3, self._scal_fpts, self.scal_upts[fout], beta=1.0)

    def negdivconf(self, fout):
        from dgfs_hip import fr
        out = self.scal_upts[fout]
        fr.negdivconf(self.nupts, self.neles, self.nvars, out, out.leaddim, self._rcpdjac_upts,
                      self._rcpdjac_upts.leaddim, self.be.stream_ptr)

    def _mul(self, A, b, out, alpha=1.0, beta=0.0):
        from dgfs_hip import fr
        m, k = A.shape
        if b.nrow != k or out.nrow != m:
            raise ValueError("operator shape mismatch")
        fr.mul(A, m, k, k, b, b.leaddim, out, out.leaddim, b.ncol, alpha, beta, self.be.stream_ptr)
