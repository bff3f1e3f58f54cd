"""Gradients of the GPU adjoint (DifferentiableEvolution, jja_adjoint) against
autograd of a CPU FP64 transcription of the scheme and against central finite
differences of the GPU forward. Stencil 3, DefaultCPR."""
import numpy as np
import pytest

import pyjjasim_amd as pa

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _circuit(inductance=0.0):
    c = pa.SquareArray(5, 4)
    rng = np.random.default_rng(3)
    c.set_resistance(0.7 + 0.6 * rng.random(c.junction_count()))
    c.set_capacitance(0.5 + rng.random(c.junction_count()))
    if inductance:
        c.set_inductance(inductance)
    return c


def _inputs(c, W, rng, dev):
    Nj, Nf = c.junction_count(), c.face_count()
    Ic = torch.tensor(np.exp(0.2 * rng.standard_normal(Nj)), dtype=torch.float64, device=dev)
    f = torch.tensor(0.1 + 0.05 * rng.standard_normal((Nf, W)), dtype=torch.float64, device=dev)
    base = c.current_base(0.0)
    Is = torch.tensor(base[:, None] * np.linspace(0.4, 1.6, W)[None, :], dtype=torch.float64,
                      device=dev)
    return Ic, f, Is


@pytest.mark.parametrize("Nt,ck,L", [(60, 16, 0.0), (60, 60, 0.0), (25, 1, 0.0), (60, 16, 0.5)])
def test_adjoint_matches_cpu_autograd(Nt, ck, L):
    from adjoint_ref import transcribe
    print(f"[adjoint] Nt={Nt} checkpoint={ck} inductance={L}", flush=True)
    c = _circuit(L)
    W, dt = 3, 0.05
    rng = np.random.default_rng(11)
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 3, Nt // 2, Nt - 1]] = True
    Ic, f, Is = _inputs(c, W, rng, "cuda")
    G = torch.tensor(rng.standard_normal((c.junction_count(), W, int(store.sum()))), dtype=torch.float64)
    evo = pa.DifferentiableEvolution(c, time_step=dt, time_step_count=Nt, problem_count=W,
                                     store_time_steps=store, checkpoint_every=ck, block_size=64)
    leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
    th = evo(*leaves)
    (th * G.cuda()).sum().backward()
    cpu = [t.detach().cpu().clone().requires_grad_(True) for t in (Ic, f, Is)]
    th_ref = transcribe(c, dt, Nt, W, store, *cpu)
    (th_ref * G).sum().backward()
    np.testing.assert_allclose(th.detach().cpu().numpy(), th_ref.detach().numpy(), rtol=1e-10, atol=1e-10)
    for name, gpu_t, cpu_t in zip(("Ic", "f", "Is"), leaves, cpu):
        want = cpu_t.grad.numpy()
        got = gpu_t.grad.cpu().numpy()
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
        print(f"  grad {name}: max rel err {err:.2e}", flush=True)
        assert err < 1e-9, name


def test_adjoint_matches_finite_differences_with_noise():
    """T > 0: the backward recomputes the same Philox noise, so the gradient of the
    noisy forward matches Richardson-extrapolated central differences of the
    forward (h = 1e-4 and 5e-5)."""
    print("[adjoint] finite differences, T=0.02", flush=True)
    c = pa.SquareArray(6, 6)
    c.set_capacitance(1.0)
    W, dt, Nt = 2, 0.05, 300
    rng = np.random.default_rng(5)
    Ic, f, Is = _inputs(c, W, rng, "cuda")
    f = f[:, 0].contiguous()                     # one flux per face, shared by the columns
    store = np.zeros(Nt, dtype=bool)
    store[-1] = True
    evo = pa.DifferentiableEvolution(c, time_step=dt, time_step_count=Nt, problem_count=W,
                                     temperature=0.02, seed=77, store_time_steps=store,
                                     checkpoint_every=64, block_size=64)

    def loss(Ic_, f_, Is_):
        th = evo(Ic_, f_, Is_)
        V = th[:, :, -1] / (Nt * dt)
        return (V ** 2).sum()
    leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
    loss(*leaves).backward()
    def central(k, i, h):
        p, m = [t.clone() for t in (Ic, f, Is)], [t.clone() for t in (Ic, f, Is)]
        p[k][i] += h
        m[k][i] -= h
        with torch.no_grad():
            return (loss(*p) - loss(*m)).item() / (2 * h)
    h = 1e-4
    for name, k, idx in (("Ic", 0, [0, 7, 30]), ("f", 1, [2, 11]), ("Is", 2, [(0, 0), (5, 1)])):
        for i in idx:
            # Richardson extrapolation of the central difference: error O(h^4)
            fd = (4.0 * central(k, i, h / 2) - central(k, i, h)) / 3.0
            ad = leaves[k].grad[i].item()
            print(f"  {name}{i}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
            # the forward's rounding, amplified over 300 noisy steps, puts a floor
            # of a few 1e-8 under these central differences (the entries here
            # differ by 1e-9 to 2e-8 whatever their size); the adjoint itself
            # matches autograd of the CPU transcription to 1e-14
            assert abs(ad - fd) <= 1e-7 * abs(fd) + 5e-8, (name, i, ad, fd)


def test_adjoint_global_face_store_bit_identical():
    """The adjoint sweep with the face store in global memory gives the same bits
    as with it in LDS."""
    print("[adjoint] global store", flush=True)
    c = _circuit()
    W, dt, Nt = 2, 0.05, 50
    rng = np.random.default_rng(21)
    Ic, f, Is = _inputs(c, W, rng, "cuda")
    store = np.zeros(Nt, dtype=bool)
    store[[20, 49]] = True
    G = torch.tensor(rng.standard_normal((c.junction_count(), W, 2)), dtype=torch.float64, device="cuda")
    grads = {}
    for fs in ("lds", "global"):
        evo = pa.DifferentiableEvolution(c, time_step=dt, time_step_count=Nt, problem_count=W,
                                         store_time_steps=store, checkpoint_every=16, block_size=64,
                                         face_store=fs)
        leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
        (evo(*leaves) * G).sum().backward()
        grads[fs] = [t.grad.cpu().numpy() for t in leaves]
    for a, b in zip(grads["lds"], grads["global"]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.timeout(600)
def test_adjoint_differentiation_workload_shape():
    """The differentiation workload's own shape: SquareArray(64,64), beta_C = 1,
    f = 0.1, W = 32 bias points 0.5..2.0, log-normal Ic (sigma 0.2), 300 steps
    with checkpoints every 200. Gradients of <G, theta_stored> in Ic, f and Is
    against autograd of the FP64 transcription (dense Cholesky solve, run with
    torch's own kernels on the device) on every junction and face, to 1e-9 of
    the largest component; the Ic gradient also against Richardson-extrapolated
    central differences (h = 1e-4, 5e-5) of the GPU forward on 50 seeded
    junctions."""
    from adjoint_ref import transcribe
    print("[adjoint] SquareArray(64,64) W=32, 300 steps", flush=True)
    c = pa.SquareArray(64, 64)
    c.set_capacitance(1.0)
    Nj, Nf = c.junction_count(), c.face_count()
    W, dt, Nt = 32, 0.05, 300
    dev = "cuda"
    Ic = torch.tensor(np.exp(0.2 * np.random.default_rng(101).standard_normal(Nj)), dtype=torch.float64, device=dev)
    f = torch.full((Nf,), 0.1, dtype=torch.float64, device=dev)
    Is = torch.tensor(c.current_base(0.0)[:, None] * np.linspace(0.5, 2.0, W)[None, :], dtype=torch.float64,
                      device=dev)
    store = np.zeros(Nt, dtype=bool)
    store[[Nt // 2, Nt - 1]] = True
    rng = np.random.default_rng(12)
    G = torch.tensor(rng.standard_normal((Nj, W, 2)), dtype=torch.float64, device=dev)
    with pa.DifferentiableEvolution(c, time_step=dt, time_step_count=Nt, problem_count=W,
                                    store_time_steps=store, checkpoint_every=200) as evo:
        leaves = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
        th = evo(*leaves)
        (th * G).sum().backward()
        ref = [t.clone().requires_grad_(True) for t in (Ic, f, Is)]
        th_ref = transcribe(c, dt, Nt, W, store, *ref, cholesky=True)
        (th_ref * G).sum().backward()
        err = (th - th_ref).abs().max().item() / th_ref.abs().max().item()
        print(f"  theta: max rel err {err:.2e}", flush=True)
        assert err < 1e-9
        for name, a, b in zip(("Ic", "f", "Is"), leaves, ref):
            e = (a.grad - b.grad).abs().max().item() / b.grad.abs().max().item()
            print(f"  grad {name}: max rel err {e:.2e}", flush=True)
            assert e < 1e-9, name
        g = leaves[0].grad.cpu().numpy()

        def loss(Ic_):
            with torch.no_grad():
                return (evo(Ic_, f, Is) * G).sum().item()

        def central(j, h):
            p, m = Ic.clone(), Ic.clone()
            p[j] += h
            m[j] -= h
            return (loss(p) - loss(m)) / (2 * h)
        idx = np.random.default_rng(303).choice(Nj, size=50, replace=False)
        h = 1e-4
        fd = np.array([(4.0 * central(j, h / 2) - central(j, h)) / 3.0 for j in idx])
    scaled = np.abs(g[idx] - fd).max() / np.abs(g).max()
    print(f"  finite differences on 50 junctions: max |adjoint - fd| / max|grad| = {scaled:.2e}", flush=True)
    assert scaled < 1e-7
