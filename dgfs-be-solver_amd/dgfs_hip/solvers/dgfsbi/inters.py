"""Boundary interfaces of the multi-species solver (mirrors
frfs/solvers/dgfsbi/inters.py:115-676).

The reference registers one comm_flux / update_bc kernel per species
(`comm_flux0`, `comm_flux1`, ...), each with the species' own boundary
distribution and flux-balance storage and otherwise the single-species
kernel arithmetic.  Here each species gets one single-species boundary
object over the same faces, holding that species' distribution
(DGFSBiMaxwellianInitCondition, n_p = 1 on walls) and accumulators; the
launches are the single-species HIP kernels.  Interior interfaces need no
species data and are the single-species ones.
"""
import torch

from dgfs_hip.solvers.dgfs import inters as I1
from dgfs_hip.solvers.dgfsbi.initcond import DGFSBiMaxwellianInitCondition


class DGFSBiBCInters:
    type = None
    single = None  # single-species class whose kernels run for every species

    def __init__(self, be, ele, faces, cfg, cfgsect, vm):
        self.be = be
        self.ele = ele
        self.cfg = cfg
        self.cfgsect = cfgsect
        self.vm = vm
        self.parts = []
        for p in range(vm.nspcs()):
            b = self.single.__new__(self.single)
            I1.DGFSBCInters.__init__(b, be, ele, faces, cfg, cfgsect)
            self._setup_species(b, p)
            self.parts.append(b)
        self.ninterfpts = self.parts[0].ninterfpts

    def _setup_species(self, b, p):
        pass

    def _species_maxwellian(self, b, p, wall):
        bc = DGFSBiMaxwellianInitCondition(self.be, self.cfg, self.vm, self.cfgsect, wall=wall)
        b._f0 = self.be.const_array(bc.get_init_vals()[p])
        u = bc.unondim()
        b._U = (float(u[0, 0]), float(u[1, 0]), float(u[2, 0]))

    @property
    def kernels(self):
        """Per-species factories under the reference's names comm_flux<p>
        and update_bc<p> (dgfsbi/inters.py: comm_flux0, comm_flux1, ...)."""
        out = {}
        for p, b in enumerate(self.parts):
            for n, f in b.kernels.items():
                out[f"{n}{p}"] = f
        return out


class DGFSBiWallDiffuseBCInters(DGFSBiBCInters):
    """inters.py:176-283: wall Maxwellian of unit number density per species,
    flux balance accumulated per species."""
    type = "dgfsbi-wall-diffuse"
    single = I1.DGFSWallDiffuseBCInters

    def _setup_species(self, b, p):
        self._species_maxwellian(b, p, wall=True)
        b._bcvals = torch.zeros(2 * max(b.ninterfpts, 1), dtype=torch.float64, device=self.be.device)


class DGFSBiWallSpecularBCInters(DGFSBiBCInters):
    """inters.py:286-421: the reflection map is species independent."""
    type = "dgfsbi-wall-specular"
    single = I1.DGFSWallSpecularBCInters

    def __init__(self, be, ele, faces, cfg, cfgsect, vm):
        self.be = be
        self.ele = ele
        self.cfg = cfg
        self.cfgsect = cfgsect
        self.vm = vm
        self.parts = [I1.DGFSWallSpecularBCInters(be, ele, faces, cfg, cfgsect) for _ in range(vm.nspcs())]
        self.ninterfpts = self.parts[0].ninterfpts


class DGFSBiWallDiffuseCylBCInters(DGFSBiBCInters):
    """inters.py:679-763, bcs/dgfsbi-wall-diffuse-cyl*.mako: the wall of
    dgfs-wall-diffuse-cyl per species, whose Maxwellian
    exp(-|c - U|^2 m_p / T) is the single-species one at temperature T / m_p
    (m_p the species' mass ratio)."""
    type = "dgfsbi-wall-diffuse-cyl"
    single = I1.DGFSWallDiffuseCylBCInters

    def __init__(self, be, ele, faces, cfg, cfgsect, vm):
        self.be = be
        self.ele = ele
        self.cfg = cfg
        self.cfgsect = cfgsect
        self.vm = vm
        self.parts = []
        for m in vm.masses():
            b = I1.DGFSWallDiffuseCylBCInters(be, ele, faces, cfg, cfgsect)
            st = b.wall_state.copy()
            st[3] /= m
            b.wall_state = st
            b._f0 = be.const_array(st.ravel())
            self.parts.append(b)
        self.ninterfpts = self.parts[0].ninterfpts


class DGFSBiInletBCInters(DGFSBiBCInters):
    """inters.py:424-495: inflow Maxwellians with the section's nden<p>."""
    type = "dgfsbi-inlet"
    single = I1.DGFSInletBCInters

    def _setup_species(self, b, p):
        self._species_maxwellian(b, p, wall=False)


class DGFSBiInletNormalShockBCInters(DGFSBiBCInters):
    """inters.py:498-560, bcs/dgfsbi-inlet-normalshock.mako: each species'
    section Maxwellian (nden<p>) on the velocities with c.n < 0."""
    type = "dgfsbi-inlet-normalshock"
    single = I1.DGFSInletNormalShockBCInters

    def _setup_species(self, b, p):
        self._species_maxwellian(b, p, wall=False)


class DGFSBiOutletBCInters(DGFSBiBCInters):
    """inters.py:563-617."""
    type = "dgfsbi-outlet"
    single = I1.DGFSOutletBCInters


class DGFSBiVanishBCInters(DGFSBiBCInters):
    """inters.py:620-676."""
    type = "dgfsbi-vanish"
    single = I1.DGFSVanishBCInters
