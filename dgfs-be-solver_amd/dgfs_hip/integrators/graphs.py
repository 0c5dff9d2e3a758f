"""HIP-graph replay of DGFS time steps.

A stepper's step() issues a few hundred launches (collision operator,
advection kernels per velocity batch, boundary updates, stage combinations;
integrators/std/steppers.py:122-225 in the reference, each launched from
Python).  GraphStepper records one step with HIP stream capture on the
backend's own stream and replays it with a single graph launch per step.

The recording holds the device pointers and scalars of that step.  A replay
is the same step because
* a DGFS step reads and writes the same registers every time: it returns the
  register it started from, which the capture checks;
* no kernel of a step depends on the time t: boundary values are functions
  of position only (the time argument of collide/rhs/updatebc is unused);
* the time step is the recorded one: a shortened last step runs eagerly.
The first step of every GraphStepper runs eagerly, whatever the stepper's
step count (a restarted stepper included), so that every workspace of the
library is allocated before the capture: no allocation may happen while
capturing.  advance_to runs the stepper's plugins after every accepted step
and counts the right-hand-side evaluations of the stepper class.
"""
import ctypes as C

from dgfs_hip import _lib

_P = C.c_void_p
_lib.register({
    "dgfs_graph_begin": [_P],
    "dgfs_graph_end": [_P, C.POINTER(C.c_void_p)],
    "dgfs_graph_launch": [_P, _P],
    "dgfs_graph_nodes": [_P, C.POINTER(C.c_long)],
    "dgfs_graph_destroy": [_P],
})


class GraphStepper:
    def __init__(self, stepper):
        be = stepper.backend
        if not getattr(be, "own_stream", False):
            raise ValueError("graph capture needs a backend with its own stream "
                             "([backend-hip] stream = own)")
        self.stepper = stepper
        self.backend = be
        self._exec = None
        self._warm = False
        self._dt = None
        self.nodes = 0

    def _capture(self, t, dt):
        st = self.backend.stream_ptr
        idx0 = self.stepper._idxcurr
        nf0 = self.stepper.nfevals
        _lib.call("dgfs_graph_begin", C.c_void_p(st))
        try:
            idx1 = self.stepper.step(t, dt)
        except Exception:
            g = C.c_void_p()
            try:
                _lib.call("dgfs_graph_end", C.c_void_p(st), C.byref(g))
            finally:
                if g.value:
                    _lib.call("dgfs_graph_destroy", g)
            raise
        self.stepper.nfevals = nf0  # recorded, not run
        g = C.c_void_p()
        _lib.call("dgfs_graph_end", C.c_void_p(st), C.byref(g))
        if idx1 != idx0:
            _lib.call("dgfs_graph_destroy", g)
            raise RuntimeError("the step changes the solution register; it cannot be replayed")
        self._exec = g
        n = C.c_long()
        _lib.call("dgfs_graph_nodes", g, C.byref(n))
        self.nodes = n.value
        # the capture recorded the step without running it
        return idx1

    def step(self, t, dt):
        """One time step: eager the first time, captured the second time (and
        then launched), replayed afterwards.  Returns the solution register."""
        s = self.stepper
        if self._exec is not None and dt != self._dt:
            raise ValueError(f"the recorded step has dt = {self._dt}; step({dt}) must run eagerly")
        if self._exec is None:
            if not self._warm:
                self._warm = True
                return s.step(t, dt)
            self._capture(t, dt)
            self._dt = dt
        _lib.call("dgfs_graph_launch", self._exec, C.c_void_p(self.backend.stream_ptr))
        s.nfevals += s._stepper_nfevals
        return s._idxcurr

    def advance_to(self, tend):
        s = self.stepper
        while s.tcurr < tend - 1e-12 * max(1.0, abs(tend)):
            dt = min(s.dt, tend - s.tcurr)
            if self._exec is not None and dt != self._dt:
                # a shortened last step differs from the recording: run it eagerly
                s._idxcurr = s.step(s.tcurr, dt)
            else:
                s._idxcurr = self.step(s.tcurr, dt)
            s.tcurr += dt
            s.nacptsteps += 1
            for pl in s.plugins:
                pl(s)

    def close(self):
        if self._exec is not None and self._exec.value:
            _lib.call("dgfs_graph_destroy", self._exec)
            self._exec = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
