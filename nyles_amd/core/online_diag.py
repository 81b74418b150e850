"""Vortex-force work diagnostic (interface of core/online_diag.py).

The pointwise work U . F_vf is not computed on the device path; ``worksum``
is reported as NaN at history times.
"""


class VFwork:
    def __init__(self, model, grid):
        self.model = model
        self.grid = grid
        self.worksum = float("nan")

    def compute(self):
        self.worksum = float("nan")
