"""Solver plug-ins on the HIP path.

DGFSResidualStdPlugin mirrors frfs/plugins/dgfsresidualstd.py:9-80: every
`nsteps` accepted steps it reports sqrt(sum ||f_n - f_{n-1}||^2 /
sum ||f_{n-1}||^2) over the full-velocity registers (previous = the state
one step earlier).  The reference copies both registers to the host; here
the previous state is a device copy and the two sums come from one HIP
reduction (dgfs_residual_norms), so only two doubles cross to the host.

DGFSDistWriterStdPlugin writes the solution files that a restart reads
(dgfs_hip.native).
"""
import math


class DGFSResidualStdPlugin:
    name = "dgfsresidualstd"
    systems = ["dgfs"]
    formulations = ["std"]

    def __init__(self, stepper, nsteps, outf=None):
        import torch
        self.stepper = stepper
        self.nsteps = int(nsteps)
        if self.nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        self.outf = outf
        be = stepper.backend
        self._work = torch.empty(1024, dtype=torch.float64, device=be.device)
        self._out2 = torch.empty(2, dtype=torch.float64, device=be.device)
        self._prev = None
        self.rows = []
        self(stepper)

    def _regs(self):
        st = self.stepper
        return [st.system.eles_scal_upts_full[st.soln_bank]]

    def _norms(self):
        """[(sum (c - p)^2, sum p^2)] per register of _regs()."""
        from dgfs_hip import fr
        st = self.stepper
        out = []
        for p, c in zip(self._prev, self._regs()):
            nupts, vsize, neles = c.ioshape
            fr.residual_norms(nupts, neles, vsize, p, c, c.leaddim, self._out2, self._work,
                              st.backend.stream_ptr)
            st.backend.synchronize()
            out.append(tuple(float(v) for v in self._out2.cpu()))
        return out

    def _row(self, norms):
        return [self.stepper.tcurr, math.sqrt(sum(a for a, _ in norms) / sum(b for _, b in norms))]

    def __call__(self, stepper=None):
        st = self.stepper
        n = st.nacptsteps
        if n % self.nsteps == 0 and n and self._prev is not None:
            row = self._row(self._norms())
            self.rows.append(row)
            if self.outf is not None:
                print(",".join(str(r) for r in row), file=self.outf)
                self.outf.flush()
            self._prev = None
        if (n + 1) % self.nsteps == 0:
            st.backend.synchronize()
            self._prev = [r.data.clone() for r in self._regs()]


class DGFSResidualBiPlugin(DGFSResidualStdPlugin):
    """Per-species residual of the multi-species system
    (plugins/dgfsresidualbi.py:9-95): row [t, r_1, ..., r_S] with r_p over
    species p's solution register (register 0 of its bank)."""
    name = "dgfsresidualbi"
    systems = ["dgfsbi"]

    def _regs(self):
        sysm = self.stepper.system
        return [sysm.species_register(p, 0) for p in range(sysm._nspcs)]

    def _row(self, norms):
        return [self.stepper.tcurr] + [math.sqrt(a / b) for a, b in norms]


class DGFSDistWriterStdPlugin:
    """Distribution writer (plugins/dgfsdistwriterstd.py:8-80 and
    dgfsdistwriterbi.py for several species): every dt_out of simulated
    time it writes the full distribution of the current solution register
    to a numbered solution file with the config, the stats (fields, prefix,
    tcurr, nsteps) and the mesh's uuid, the input of a restart.  Writes the
    initial state unless restarting."""
    name = "dgfsdistwriterstd"
    systems = ["dgfs", "dgfsbi"]
    formulations = ["std"]

    def __init__(self, stepper, dt_out, basedir, basename, isrestart=False, container="auto"):
        from dgfs_hip.native import NativeWriter
        self.stepper = stepper
        self.dt_out = float(dt_out)
        self.writer = NativeWriter(basedir, basename, prefix="soln", isrestart=isrestart,
                                   container=container)
        # output times are met to within a few ulps of the accumulated time
        # (the reference's plugins use 5 dt_min, base.py)
        self.tol = 5e-12 * max(1.0, abs(stepper.tcurr) + self.dt_out)
        self.tout_next = stepper.tcurr
        self.paths = []
        if not isrestart:
            self(stepper)
        else:
            self.tout_next += self.dt_out

    def __call__(self, stepper=None):
        from dgfs_hip.inifile import Inifile
        st = self.stepper
        if abs(self.tout_next - st.tcurr) > self.tol:
            return
        sysm = st.system
        stats = Inifile()
        stats.set("data", "fields", sysm.dist_fields())
        stats.set("data", "prefix", "soln")
        stats.set("solver-time-integrator", "tcurr", repr(float(st.tcurr)))
        stats.set("solver-time-integrator", "nsteps", st.nacptsteps)
        metadata = dict(config=sysm.cfg.tostr(), stats=stats.tostr(), mesh_uuid=sysm.mesh.uuid)
        self.paths.append(self.writer.write(sysm.soln_arrays(st.soln_bank), metadata, st.tcurr))
        self.tout_next = st.tcurr + self.dt_out
