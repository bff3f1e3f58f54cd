"""Host-side layout of the device inputs: every broadcastable input shape the
reference accepts (time_evolution.py:119-130) maps to a compact [t][w][i] array
whose strides reproduce every element of the broadcast (n, W, Nt) view."""
import numpy as np
import pytest

from pyjjasim_amd.engine import source_layout

N, W, NT = 5, 3, 7


@pytest.mark.parametrize("shape", [(), (1, 1, 1), (N, 1, 1), (N, W, 1), (1, W, 1), (N, 1, NT),
                                   (N, W, NT), (1, 1, NT), (W, 1), (1, NT), (NT,)])
def test_source_layout_reproduces_broadcast(shape):
    rng = np.random.default_rng(len(shape) * 10 + sum(shape))
    value = rng.standard_normal(shape) if shape else 0.25
    full = np.broadcast_to(np.asarray(value, dtype=np.float64), (N, W, NT))
    flat, (s_i, s_w, s_t) = source_layout(value, N, W, NT)
    i, w, t = np.meshgrid(np.arange(N), np.arange(W), np.arange(NT), indexing="ij")
    np.testing.assert_array_equal(flat[i * s_i + w * s_w + t * s_t], full)
    assert flat.size == np.prod([d for d, s in zip((N, W, NT), (s_i, s_w, s_t)) if s] or [1])
    if s_i:
        assert s_i == 1, "element index must be the fastest axis"


def test_auto_parts_choice(monkeypatch):
    """Parts per column: 1 for an LDS face store or another stencil; otherwise
    the largest power of two (at most 16) whose workgroups all fit one per
    compute unit and keep at least 24k junctions each."""
    from types import SimpleNamespace
    from pyjjasim_amd import engine
    monkeypatch.setattr(engine, "device_compute_units", lambda: 256)
    big = SimpleNamespace(global_store=True, block_size=1024, Nj=523264)
    assert engine.auto_parts(big, 16, 3) == 16          # 16 columns x 16 parts = 256
    assert engine.auto_parts(big, 64, 3) == 4           # 64 x 4 = 256
    assert engine.auto_parts(big, 256, 3) == 1          # no room left
    assert engine.auto_parts(big, 16, 4) == 1           # split columns: stencil 3 only
    mid = SimpleNamespace(global_store=True, block_size=1024, Nj=70000)
    assert engine.auto_parts(mid, 1, 3) == 2            # >= 24k junctions per part
    lds = SimpleNamespace(global_store=False, block_size=1024, Nj=523264)
    assert engine.auto_parts(lds, 1, 3) == 1
    odd = SimpleNamespace(global_store=True, block_size=128, Nj=523264)
    assert engine.auto_parts(odd, 1, 3) == 1            # no split kernel for 128 threads
