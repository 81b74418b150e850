"""Time schemes (interface of core/timescheme.py).

``Timescheme.forward`` is one of ``EulerForward``, ``LFAM3`` or ``RK3_SSP``
(core/timescheme.py:62-99); each runs the whole step on the GPU through
the model's device context: the stage updates are fused into the
right-hand-side kernel and every stage is followed by the model's
diagnose_var (projection), exactly as the reference sequences them
(core/timescheme.py:113-221).  A forced model (user forcing on the host
views of the tendencies) runs each stage as rhs + forcing (the model's
``rhs``), then the stage update from the stored tendencies, then
diagnose_var: the same arithmetic, in separate device calls.
"""


class Timescheme(object):
    def __init__(self, param, state):
        self.prognostic_scalars = state.get_prognostic_scalars()
        timestepping = param["timestepping"]
        schemes = {"EF": self.EulerForward, "LFAM3": self.LFAM3, "RK3_SSP": self.RK3_SSP}
        try:
            self.forward = schemes[timestepping]
        except KeyError:
            raise ValueError("unknown time scheme " + repr(timestepping))
        self.timestepping = timestepping
        self.dstate = state.duplicate_prognostic_variables()
        if timestepping == "LFAM3":
            self.stateb = state.duplicate_prognostic_variables()
            self.state = state.duplicate_prognostic_variables()
            self.first = True
        if timestepping == "RK3_SSP":
            self.ds0 = self.dstate
            self.ds1 = state.duplicate_prognostic_variables()
            self.ds2 = state.duplicate_prognostic_variables()
        self.device = None
        self.stats = {}

    def set(self, rhs, diagnose_var):
        self.rhs = rhs
        self.diagnose_var = diagnose_var

    def bind(self, device_model, model=None):
        """Run the steps on this DeviceModel (model: the LES owning it, whose
        forcing, if any, makes the stages run unfused)."""
        self.device = device_model
        self.model = model

    def _need_device(self):
        if self.device is None:
            raise RuntimeError("the time scheme is not bound to a device model")

    def _unfused(self):
        return self.model is not None and bool(getattr(self.model, "forced", False))

    def _stage(self, state, t, dstate, last, mode, c, c2=0.0):
        """rhs (+ forcing) into dstate's slot, the stage update of `mode`,
        diagnose_var; returns the projection's V-cycles"""
        self.rhs(state, t, dstate, last=last)
        self.model.push()
        self.device.stage_update(mode, c, c2)
        self.model.mark_device_newer()
        self.diagnose_var(state)
        return self.device.last_stats.as_dict()

    def _sum(self, parts):
        out = dict(parts[-1])
        out["nite"] = sum(p["nite"] for p in parts)
        return out

    def EulerForward(self, state, t, dt, **kwargs):
        self._need_device()
        if self._unfused():
            self.stats = self._stage(state, t, self.dstate, True, 1, dt)
            return
        self.stats = self.device.step_ef(dt)

    def LFAM3(self, state, t, dt, **kwargs):
        self._need_device()
        if self._unfused():
            if self.first:
                self.stats = self._stage(state, t, self.dstate, False, 6, dt)
            else:
                p = self._stage(state, t, self.dstate, False, 4, 0.0, 2. * dt)
                c = self._stage(state, t + dt * .5, self.dstate, True, 5, dt)
                self.stats = self._sum([p, c])
            self.first = False
            return
        self.stats = self.device.step_lfam3(dt)
        self.first = False

    def RK3_SSP(self, state, t, dt, **kwargs):
        self._need_device()
        if self._unfused():
            parts = [self._stage(state, t, self.ds0, False, 1, dt),
                     self._stage(state, t + dt, self.ds1, False, 2, dt / 4.),
                     self._stage(state, t + dt * 0.5, self.ds2, True, 3, dt / 12.)]
            self.stats = self._sum(parts)
            return
        self.stats = self.device.step_rk3(dt)
