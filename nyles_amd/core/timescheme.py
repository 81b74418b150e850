"""Time schemes (interface of core/timescheme.py).

``Timescheme.forward`` is one of ``EulerForward``, ``LFAM3`` or ``RK3_SSP``
(core/timescheme.py:62-99); each runs the whole step on the GPU through
the model's device context: the stage updates are fused into the
right-hand-side kernel and every stage is followed by the model's
diagnose_var (projection), exactly as the reference sequences them
(core/timescheme.py:113-221).
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

    def bind(self, device_model):
        """Run the steps on this DeviceModel."""
        self.device = device_model

    def _need_device(self):
        if self.device is None:
            raise RuntimeError("the time scheme is not bound to a device model")

    def EulerForward(self, state, t, dt, **kwargs):
        self._need_device()
        self.stats = self.device.step_ef(dt)

    def LFAM3(self, state, t, dt, **kwargs):
        self._need_device()
        self.stats = self.device.step_lfam3(dt)
        self.first = False

    def RK3_SSP(self, state, t, dt, **kwargs):
        self._need_device()
        self.stats = self.device.step_rk3(dt)
