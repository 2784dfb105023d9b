"""Host-side logic without a GPU: the CFL time-step controller."""
import pytest

from mydemcfd_amd.cfdtime import EPS, CFDTime


def test_cfdtime_first_steps_capped_by_initdt():
    t = CFDTime(initdt=1e-3, max_dt=1.0, out_freq_time=10.0, end_time=100.0, cfl_thresh=0.3)
    # small CFL would allow a large dt: the first five steps keep initdt
    for _ in range(5):
        assert t.update_time(0.01) == 1e-3
    # a large CFL shrinks dt below initdt even in the first steps
    t = CFDTime(initdt=1e-3, max_dt=1.0, out_freq_time=10.0, end_time=100.0, cfl_thresh=0.3)
    assert t.update_time(0.6) == pytest.approx(1e-3 * 0.3 / (0.6 + EPS), rel=1e-15)


def test_cfdtime_growth_clamps_and_output_times():
    t = CFDTime(initdt=0.1, max_dt=0.25, out_freq_time=1.0, end_time=3.0, cfl_thresh=0.5)
    for _ in range(5):
        t.update_time(0.5)  # cfl == thresh: dt stays initdt
    assert t.current_steps == 5 and t.current_time == pytest.approx(0.5)
    dt = t.update_time(0.25)  # doubles to 0.2, ends at 0.7
    assert dt == pytest.approx(0.2) and not t.is_out_step
    dt = t.update_time(0.1)  # 0.2*5 = 1.0 -> cut to the 0.3 left to the output, then to max_dt
    assert dt == 0.25 and not t.is_out_step
    dt = t.update_time(0.01)  # large growth -> cut to the 0.05 left: an output step
    assert dt == pytest.approx(0.05) and t.is_out_step and t.next_out == pytest.approx(2.0)
    assert t.current_time == pytest.approx(1.0, abs=1e-9)
    while not t.done():
        t.update_time(0.5)
    assert t.current_time == pytest.approx(3.0, abs=1e-12) or t.current_time > 3.0
