"""SeparableSource stands for base[:, :, None] * amplitude[None, :, :] with the
reference's (N, W, Nt) broadcasting (time_evolution.py:119-130)."""
import numpy as np
import pytest

import pyjjasim_amd as pa
from pyjjasim_amd.engine import source_layout


def test_values_equal_materialised_product():
    rng = np.random.default_rng(0)
    base, amp = rng.standard_normal((7, 3)), rng.standard_normal((3, 11))
    s = pa.SeparableSource(base, amp)
    full = base[:, :, None] * amp[None, :, :]
    assert s.shape == (7, 3, 11)
    np.testing.assert_array_equal(np.asarray(s), full)
    np.testing.assert_array_equal(s[:, :, 4], full[:, :, 4])
    np.testing.assert_array_equal(s[2:5, 1:, 3:8], full[2:5, 1:, 3:8])


def test_broadcast_forms():
    base, amp = np.arange(1.0, 6.0), 0.5 + np.sin(np.arange(9.0))
    s = pa.SeparableSource(base, amp)  # (N,) x (Nt,) -> (N, 1, Nt)
    assert s.shape == (5, 1, 9)
    np.testing.assert_array_equal(np.asarray(s)[:, 0, :], np.outer(base, amp))
    s2 = pa.SeparableSource(base, np.stack([amp, 2 * amp]))  # W from the amplitude
    assert s2.shape == (5, 2, 9)
    np.testing.assert_array_equal(np.asarray(s2)[:, 1, :], np.outer(base, 2 * amp))
    with pytest.raises(ValueError):
        pa.SeparableSource(np.ones((5, 3)), np.ones((2, 9)))


def test_problem_accepts_separable_source():
    c = pa.SquareArray(4, 3)
    Nt = 20
    drive = pa.SeparableSource(c.current_base(0.0), np.linspace(0.0, 1.0, Nt))
    p = pa.TimeEvolutionProblem(c, time_step=0.1, time_step_count=Nt, current_sources=drive)
    assert p.get_problem_count() == 1
    np.testing.assert_array_equal(np.asarray(p.current_sources), np.asarray(drive))


def test_device_layout_strides():
    """The [t][w][i] layout and strides the device reads: a broadcast axis gets
    stride 0, so element (i, w, t) sits at i*s_i + w*s_w + t*s_t."""
    rng = np.random.default_rng(1)
    for shape, W, Nt in [((6, 1, 1), 4, 10), ((6, 4, 1), 4, 10), ((6, 1, 10), 4, 10), ((6, 4, 10), 4, 10)]:
        v = rng.standard_normal(shape)
        arr, (s_i, s_w, s_t) = source_layout(v, 6, W, Nt)
        flat = np.ravel(arr)
        full = np.broadcast_to(v, (6, W, Nt))
        for i, w, t in [(0, 0, 0), (5, 3, 9), (2, 1, 7)]:
            assert flat[i * s_i + w * s_w + t * s_t] == full[i, w, t]
