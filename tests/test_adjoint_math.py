"""CPU checks of the adjoint recurrence the GPU kernel implements (jja_adjoint):
the torch transcription used as the gradient reference reproduces the oracle's
trajectory, and the recurrence
    mu_k = M lambda_k,  M = D A^T K^-1 A D - D,
    lambda_{k-1} += (2 Ic cos a_k + c1) mu_k,  lambda_{k-2} += (c2 - Ic cos a_k) mu_k,
    dIc += sin(a_k) mu_k,  dIs -= mu_k,  df -= 2 pi K^-1 A D lambda_k
run in numpy gives autograd's gradients of the transcription."""
import numpy as np
import pytest

import pyjjasim_amd as pa
from oracle import te_oracle

torch = pytest.importorskip("torch")


def _setup():
    c = pa.SquareArray(4, 5)
    rng = np.random.default_rng(9)
    c.set_resistance(0.8 + 0.4 * rng.random(c.junction_count()))
    c.set_capacitance(0.3 + rng.random(c.junction_count()))
    Nj, Nf, W = c.junction_count(), c.face_count(), 2
    Ic = np.exp(0.2 * rng.standard_normal(Nj))
    f = 0.1 + 0.05 * rng.standard_normal((Nf, W))
    Is = c.current_base(0.0)[:, None] * np.array([0.6, 1.3])[None, :]
    return c, rng, Ic, f, Is, W


def test_transcription_matches_oracle():
    from adjoint_ref import transcribe
    c, rng, Ic, f, Is, W = _setup()
    c.set_critical_current(Ic)
    dt, Nt = 0.05, 40
    store = np.ones(Nt, dtype=bool)
    th = transcribe(c, dt, Nt, W, store, torch.tensor(Ic), torch.tensor(f), torch.tensor(Is)).numpy()
    Nj, Nf = c.junction_count(), c.face_count()
    inputs = te_oracle.Inputs(Nj, Nf, Nt, f[:, :, None], Is[:, :, None], 0.0, 0.0)
    zero = np.zeros((Nj, W))
    th_o, _ = te_oracle.evolve_core(c, dt, Nt, inputs, store, np.zeros(Nt, dtype=bool), [zero, zero])
    np.testing.assert_allclose(th, th_o[:, :, 2:], rtol=1e-10, atol=1e-11)


def test_adjoint_recurrence_matches_autograd():
    from adjoint_ref import transcribe
    from pyjjasim_amd.solver_plan import STENCILS, face_system_matrix
    c, rng, Ic, f, Is, W = _setup()
    dt, Nt = 0.05, 30
    store = np.zeros(Nt, dtype=bool)
    store[[9, 20, 29]] = True
    G = rng.standard_normal((c.junction_count(), W, 3))
    leaves = [torch.tensor(v, requires_grad=True) for v in (Ic, f, Is)]
    th = transcribe(c, dt, Nt, W, store, *leaves)
    (th * torch.tensor(G)).sum().backward()
    # forward history in numpy
    with torch.no_grad():
        hist = transcribe(c, dt, Nt, W, np.ones(Nt, dtype=bool), *[torch.tensor(v) for v in (Ic, f, Is)]).numpy()
    Nj = c.junction_count()
    thm = lambda m: np.zeros((Nj, W)) if m < 0 else hist[:, :, m]
    A = c.get_cycle_matrix().toarray()
    Rv, Cv = 1.0 / (dt * c._R()), c._C() / dt ** 2
    S1, S2 = STENCILS[3]
    c0, c1, c2 = (S1[k] * Rv + S2[k] * Cv for k in range(3))
    K = np.asarray(face_system_matrix(c, dt).todense())
    Kinv_A_D = np.linalg.solve(K, A / c0[None, :])
    M = (A.T @ Kinv_A_D) / c0[:, None] - np.diag(1.0 / c0)
    g = np.zeros((Nj, W, Nt))
    g[:, :, store] = G
    lam = {Nt - 1: g[:, :, Nt - 1].copy(), Nt - 2: np.zeros((Nj, W))}
    gIc, gIs, gf = np.zeros(Nj), np.zeros((Nj, W)), np.zeros(f.shape)
    for k in range(Nt - 1, -1, -1):
        lk = lam.pop(k)
        mu = M @ lk
        a_k = 2 * thm(k - 1) - thm(k - 2)
        gIc += (np.sin(a_k) * mu).sum(1)
        gIs -= mu
        gf -= 2 * np.pi * (Kinv_A_D @ lk)
        lam[k - 1] = lam.get(k - 1, 0) + (2 * Ic[:, None] * np.cos(a_k) + c1[:, None]) * mu + \
            (g[:, :, k - 1] if k >= 1 else 0)
        lam[k - 2] = (c2[:, None] - Ic[:, None] * np.cos(a_k)) * mu
    for name, got, leaf in (("Ic", gIc, leaves[0]), ("f", gf, leaves[1]), ("Is", gIs, leaves[2])):
        want = leaf.grad.numpy()
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max(), err_msg=name)
