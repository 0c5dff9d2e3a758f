"""DGFS time steppers (mirrors frfs/integrators/std/steppers.py:20-225 and the
register plumbing of integrators/base.py:152-214).

The sequence of collide / updatebc / rhs / axnpby / register copies of each
stage is the reference's; every operation is a HIP launch on the backend
stream.  The stage kernels come from the backend's provider walk
(backend.kernel('axnpby' | 'copy_to_reg' | 'copy_from_reg', ...)), as the
reference's integrators get them (integrators/base.py:152-214).
"""


class BaseDGFSStepper:
    stepper_name = None

    def __init__(self, backend, system, dt, tstart=0.0):
        self.backend = backend
        self.system = system
        self.dt = dt
        self.tcurr = tstart
        self.nacptsteps = 0
        self._idxcurr = 0
        self._regidx = list(range(self._stepper_nregs))
        self._axk = {}
        self._cpk = {}
        self.nfevals = 0
        # called as plugin(stepper) after every accepted step of advance_to
        # (the reference integrator's completed-step handlers)
        self.plugins = []

    # register plumbing ----------------------------------------------------
    def _add_dgfs(self, *args):
        """_add_dgfs(a0, r0, a1, r1, ...): x_r0 = a0 x_r0 + a1 x_r1 + ... on the
        batch registers (integrators/base.py:194-214)."""
        idx = tuple(args[1::2])
        if idx not in self._axk:
            regs = self.system.ele.scal_upts
            self._axk[idx] = self.backend.kernel("axnpby", *[regs[i] for i in idx])
        self._axk[idx].run(None, *args[::2])

    def _copy_to_reg_dgfs(self, inp, out, sidx, eidx, rsidx, reidx):
        key = ("to", inp, out)
        if key not in self._cpk:
            self._cpk[key] = self.backend.kernel("copy_to_reg", self.system.eles_scal_upts_full[inp],
                                                 self.system.ele.scal_upts[out])
        self._cpk[key].run(None, sidx, eidx, rsidx, reidx)

    def _copy_from_reg_dgfs(self, inp, out, rsidx, reidx, sidx, eidx):
        key = ("from", inp, out)
        if key not in self._cpk:
            self._cpk[key] = self.backend.kernel("copy_from_reg", self.system.ele.scal_upts[inp],
                                                 self.system.eles_scal_upts_full[out])
        self._cpk[key].run(None, rsidx, reidx, sidx, eidx)

    def advance_to(self, tend):
        while self.tcurr < tend - 1e-12 * max(1.0, abs(tend)):
            dt = min(self.dt, tend - self.tcurr)
            self._idxcurr = self.step(self.tcurr, dt)
            self.tcurr += dt
            self.nacptsteps += 1
            for pl in self.plugins:
                pl(self)

    @property
    def soln_bank(self):
        return self._idxcurr

    def restart(self, initsoln):
        """Resume the clock of a solution file: tcurr and the accepted step
        count (integrators/base.py:35-38)."""
        from dgfs_hip.native import restart_stats
        self.tcurr, self.nacptsteps = restart_stats(initsoln)


class StdEulerDGFSStepper(BaseDGFSStepper):
    stepper_name = "dgfs-euler"
    _stepper_nregs = 2
    _stepper_order = 1
    _stepper_nfevals = 1     # right-hand-side evaluations per step

    def step(self, t, dt):
        add, rhs = self._add_dgfs, self.system.rhs
        collide, updatebc = self.system.collide, self.system.updatebc
        copyto, copyfrm = self._copy_to_reg_dgfs, self._copy_from_reg_dgfs
        vm = self.system.vm
        NvBS = vm.NvBatchSize()
        r0, r1 = self._regidx
        if r0 != self._idxcurr:
            r0, r1 = r1, r0
        collide(t, r0, r1)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
            updatebc(t, r0, r1, batch, sidx)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
            rhs(t, r0, r1, batch, sidx)
            add(1.0, r0, dt, r1)
            copyto(r1, r1, sidx, eidx - 1, 0, NvBS - 1)
            add(1.0, r0, dt, r1)
            copyfrm(r0, r0, 0, NvBS - 1, sidx, eidx - 1)
        self.nfevals += self._stepper_nfevals
        return r0


class StdTVDRK2DGFSStepper(BaseDGFSStepper):
    stepper_name = "dgfs-tvd-rk2"
    _stepper_nregs = 3
    _stepper_order = 2
    _stepper_nfevals = 2     # right-hand-side evaluations per step

    def step(self, t, dt):
        add, rhs = self._add_dgfs, self.system.rhs
        collide, updatebc = self.system.collide, self.system.updatebc
        copyto, copyfrm = self._copy_to_reg_dgfs, self._copy_from_reg_dgfs
        vm = self.system.vm
        NvBS = vm.NvBatchSize()
        r0, r1, r2 = self._regidx
        if r0 != self._idxcurr:
            r0, r1 = r1, r0
        # stage 1: r1 = r0 + dt (adv(r0) + Q(r0))
        collide(t, r0, r1)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
            updatebc(t, r0, r1, batch, sidx)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
            rhs(t, r0, r1, batch, sidx)
            add(1.0, r0, dt, r1)
            copyto(r1, r1, sidx, eidx - 1, 0, NvBS - 1)
            add(dt, r1, 1.0, r0)
            copyfrm(r1, r1, 0, NvBS - 1, sidx, eidx - 1)
        # stage 2: r0 = 0.5 r0 + 0.5 r1 + 0.5 dt (adv(r1) + Q(r1))
        collide(t + dt, r1, r2)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r1, r1, sidx, eidx - 1, 0, NvBS - 1)
            updatebc(t, r1, r2, batch, sidx)
        for batch in range(vm.NvBatches()):
            sidx, eidx = batch * NvBS, (batch + 1) * NvBS
            copyto(r2, r2, sidx, eidx - 1, 0, NvBS - 1)
            copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
            add(0.5, r0, 0.5 * dt, r2)
            copyto(r1, r1, sidx, eidx - 1, 0, NvBS - 1)
            rhs(t + dt, r1, r2, batch, sidx)
            add(1.0, r0, 0.5, r1, 0.5 * dt, r2)
            copyfrm(r0, r0, 0, NvBS - 1, sidx, eidx - 1)
        self.nfevals += self._stepper_nfevals
        return r0


class BaseDGFSBiStepper(BaseDGFSStepper):
    """Multi-species steppers (steppers.py:229-476): registers are banked
    species-major, nregs_orig per species; the solution of every species
    stays in its register 0 (the reference raises otherwise)."""
    _stepper_nregs_orig = None

    def __init__(self, backend, system, dt, tstart=0.0):
        self._stepper_nregs = self._stepper_nregs_orig * system._nspcs
        super().__init__(backend, system, dt, tstart)

    def _check_current(self):
        if self._idxcurr != 0:
            raise ValueError("Something wrong")

    def _collide(self, t, rin, rout):
        n0 = self._stepper_nregs_orig
        self.system.collide(t, list(self._regidx[rin::n0]) + list(self._regidx[rout::n0]))


class StdEulerDGFSBiStepper(BaseDGFSBiStepper):
    stepper_name = "dgfsbi-euler"
    _stepper_nregs_orig = 2
    _stepper_order = 1
    _stepper_nfevals = 1     # right-hand-side evaluations per step

    def step(self, t, dt):
        add, rhs = self._add_dgfs, self.system.rhs
        updatebc = self.system.updatebc
        copyto, copyfrm = self._copy_to_reg_dgfs, self._copy_from_reg_dgfs
        vm = self.system.vm
        NvBS, n0 = vm.NvBatchSize(), self._stepper_nregs_orig
        self._check_current()
        self._collide(t, 0, 1)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
                updatebc(t, str(p), r0, r0 + 1, batch, sidx)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
                rhs(t, str(p), r0, r0 + 1, batch, sidx)
                add(1.0, r0, dt, r0 + 1)
                copyto(r0 + 1, r0 + 1, sidx, eidx - 1, 0, NvBS - 1)
                add(1.0, r0, dt, r0 + 1)
                copyfrm(r0, r0, 0, NvBS - 1, sidx, eidx - 1)
        self.nfevals += self._stepper_nfevals
        return 0


class StdTVDRK2DGFSBiStepper(BaseDGFSBiStepper):
    stepper_name = "dgfsbi-tvd-rk2"
    _stepper_nregs_orig = 3
    _stepper_order = 2
    _stepper_nfevals = 2     # right-hand-side evaluations per step

    def step(self, t, dt):
        add, rhs = self._add_dgfs, self.system.rhs
        updatebc = self.system.updatebc
        copyto, copyfrm = self._copy_to_reg_dgfs, self._copy_from_reg_dgfs
        vm = self.system.vm
        NvBS, n0 = vm.NvBatchSize(), self._stepper_nregs_orig
        self._check_current()
        # stage 1: r1 = r0 + dt (adv(r0) + Q(r0)) for every species
        self._collide(t, 0, 1)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
                updatebc(t, str(p), r0, r0 + 1, batch, sidx)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
                rhs(t, str(p), r0, r0 + 1, batch, sidx)
                add(1.0, r0, dt, r0 + 1)
                copyto(r0 + 1, r0 + 1, sidx, eidx - 1, 0, NvBS - 1)
                add(dt, r0 + 1, 1.0, r0)
                copyfrm(r0 + 1, r0 + 1, 0, NvBS - 1, sidx, eidx - 1)
        # stage 2: r0 = 0.5 r0 + 0.5 r1 + 0.5 dt (adv(r1) + Q(r1))
        self._collide(t, 1, 2)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0 + 1, r0 + 1, sidx, eidx - 1, 0, NvBS - 1)
                updatebc(t, str(p), r0 + 1, r0 + 2, batch, sidx)
        for p in range(vm.nspcs()):
            r0 = n0 * p
            for batch in range(vm.NvBatches()):
                sidx, eidx = batch * NvBS, (batch + 1) * NvBS
                copyto(r0 + 2, r0 + 2, sidx, eidx - 1, 0, NvBS - 1)
                copyto(r0, r0, sidx, eidx - 1, 0, NvBS - 1)
                add(0.5, r0, 0.5 * dt, r0 + 2)
                copyto(r0 + 1, r0 + 1, sidx, eidx - 1, 0, NvBS - 1)
                rhs(t + dt, str(p), r0 + 1, r0 + 2, batch, sidx)
                add(1.0, r0, 0.5, r0 + 1, 0.5 * dt, r0 + 2)
                copyfrm(r0, r0, 0, NvBS - 1, sidx, eidx - 1)
        self.nfevals += self._stepper_nfevals
        return 0


STEPPERS = {c.stepper_name: c for c in (StdEulerDGFSStepper, StdTVDRK2DGFSStepper,
                                        StdEulerDGFSBiStepper, StdTVDRK2DGFSBiStepper)}


def get_stepper(name, backend, system, dt, tstart=0.0):
    if name not in STEPPERS:
        raise KeyError(f"unknown stepper {name!r}: {sorted(STEPPERS)}")
    return STEPPERS[name](backend, system, dt, tstart)
