"""CFL time-step controller of the reference loop (CFD/CFDTime.cpp:22-48).

The first five steps take min(initdt, dt*cfl_thresh/(cfl+EPS)); later steps
take dt*cfl_thresh/(cfl+EPS), cut so that the step ends on the next output
time and to max_dt; a step that lands within 1e-8 of the output time sets
is_out_step and advances the next output time. `cfl` is the larger of the
convective and viscous CFL numbers of the current dt (smac_calc_cfl).
"""
from __future__ import annotations

EPS = 1.0e-12  # CFD/hardCodedParameters.h:7


class CFDTime:
    def __init__(self, initdt, max_dt, out_freq_time, end_time, cfl_thresh):
        self.dt = initdt
        self.initdt = initdt
        self.max_dt = max_dt
        self.out_freq_time = out_freq_time
        self.next_out = out_freq_time
        self.end_time = end_time
        self.cfl_thresh = cfl_thresh
        self.current_time = 0.0
        self.current_steps = 0
        self.is_out_step = False

    def update_time(self, cfl):
        """Advance to the next step from the CFL number of the current dt."""
        if self.current_steps < 5:
            dt_tmp = self.dt * self.cfl_thresh / (cfl + EPS)
            self.dt = self.initdt if self.initdt < dt_tmp else dt_tmp
        else:
            self.dt = self.dt * self.cfl_thresh / (cfl + EPS)
            till_out = self.next_out - self.current_time
            if self.dt >= till_out:
                self.dt = till_out
            if self.dt > self.max_dt:
                self.dt = self.max_dt
            if abs(self.current_time + self.dt - self.next_out) <= 1e-8:
                self.is_out_step = True
                self.next_out += self.out_freq_time
        self.current_time += self.dt
        self.current_steps += 1
        return self.dt

    def done(self):
        return self.current_time >= self.end_time
