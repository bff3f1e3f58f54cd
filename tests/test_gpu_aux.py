"""Checkpoint/resume and invalid values on the device path.

* Resume (time_evolution.py:71-75, 144-149): a run continued from its own last
  two phases as config_at_minus_1/2 reproduces the uninterrupted run bit for
  bit (T = 0: nothing depends on the absolute step).
* NaN input: a NaN critical current propagates to the phases it reaches (as
  numpy would) without a fault, a hang or a status error, and the next run on
  the device is exact again."""
import numpy as np
import pytest

import pyjjasim_amd as pa
from oracle import te_oracle as to

pytestmark = pytest.mark.gpu


def _circuit():
    c = pa.SquareArray(12, 9)
    c.set_capacitance(0.7)
    c.set_inductance(0.3)
    return c


def test_gpu_resume_from_last_two_phases_is_bit_identical():
    print("[aux] resume", flush=True)
    c = _circuit()
    Is = c.current_base(0.0)[:, None, None] * np.array([0.6, 1.4])[None, :, None]
    kw = dict(time_step=0.05, external_flux=0.15, current_sources=Is, store_current=False, store_voltage=False)
    full = pa.TimeEvolutionProblem(c, time_step_count=400, store_time_steps=[198, 199, 399], **kw).compute()
    first = pa.TimeEvolutionProblem(c, time_step_count=200, store_time_steps=[198, 199], **kw).compute()
    th1 = first.get_theta()
    second = pa.TimeEvolutionProblem(c, time_step_count=200, store_time_steps=[199], config_at_minus_1=th1[:, :, 1],
                                     config_at_minus_2=th1[:, :, 0], **kw).compute()
    np.testing.assert_array_equal(first.get_theta(), full.get_theta()[:, :, :2])
    np.testing.assert_array_equal(second.get_theta()[:, :, 0], full.get_theta()[:, :, 2])


def test_gpu_nan_input_propagates_without_fault():
    print("[aux] NaN critical current", flush=True)
    c = _circuit()
    Ic = np.ones(c.junction_count())
    Ic[17] = np.nan
    c.set_critical_current(Ic)
    Is = 0.8 * c.current_base(0.0)[:, None, None]
    res = pa.TimeEvolutionProblem(c, time_step=0.05, time_step_count=50, external_flux=0.1, current_sources=Is,
                                  store_time_steps=[0, 49]).compute()
    th = res.get_theta()
    assert np.isnan(th[17, 0, 0]) and np.isnan(th[:, 0, 1]).any()
    ref = to.evolve(c, dt=0.05, Nt=50, f=0.1, Is=Is, store_time_steps=[0, 49])
    np.testing.assert_array_equal(np.isnan(th), np.isnan(ref["theta"]))
    good = _circuit()
    res2 = pa.TimeEvolutionProblem(good, time_step=0.05, time_step_count=50, external_flux=0.1, current_sources=Is,
                                   store_time_steps=[49]).compute()
    ref2 = to.evolve(good, dt=0.05, Nt=50, f=0.1, Is=Is, store_time_steps=[49])
    assert np.abs(res2.get_theta() - ref2["theta"]).max() <= 1e-9 * np.abs(ref2["theta"]).max()
