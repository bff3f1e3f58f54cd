"""FP64 torch transcription of the stencil-3 step (time_evolution.py:523-580)
for gradients by autograd: test infrastructure for the adjoint kernel. Dense K
(solved by LU each step, or through its dense Cholesky factor computed once),
T = 0, constant flux and bias; it runs on whatever device its inputs are on
(torch's own FP64 kernels, none of this package's)."""
import numpy as np
import torch

from pyjjasim_amd.solver_plan import STENCILS, face_system_matrix


def transcribe(circuit, dt, Nt, W, store, Ic, f, Is, th_m1=None, th_m2=None, cholesky=False):
    """theta at the stored steps, (Nj, W, n_stored), differentiable in Ic (Nj,),
    f (Nf,) or (Nf, W) and Is (Nj,) or (Nj, W), on the device of Ic."""
    dev = Ic.device
    A = torch.from_numpy(circuit.get_cycle_matrix().toarray()).double().to(dev)
    Nj = A.shape[1]
    Rv = 1.0 / (dt * circuit._R())
    Cv = circuit._C() / dt ** 2
    S1, S2 = STENCILS[3]
    c0 = torch.from_numpy(S1[0] * Rv + S2[0] * Cv)[:, None].to(dev)
    c1 = torch.from_numpy(S1[1] * Rv + S2[1] * Cv)[:, None].to(dev)
    c2 = torch.from_numpy(S1[2] * Rv + S2[2] * Cv)[:, None].to(dev)
    K = torch.from_numpy(np.asarray(face_system_matrix(circuit, dt).todense())).double().to(dev)
    Lk = torch.linalg.cholesky(K) if cholesky else None
    fb = f[:, None] if f.dim() == 1 else f
    Isb = Is[:, None] if Is.dim() == 1 else Is
    th1 = torch.zeros((Nj, W), dtype=torch.float64, device=dev) if th_m1 is None else th_m1
    th2 = th1.clone() if th_m2 is None else th_m2
    outs = []
    for i in range(Nt):
        X = Ic[:, None] * torch.sin(2 * th1 - th2) + c1 * th1 + c2 * th2
        x = -Isb + X
        b = A @ (x / c0) - 2 * np.pi * fb
        J = torch.cholesky_solve(b, Lk) if cholesky else torch.linalg.solve(K, b)
        th = (A.T @ J - x) / c0
        if store[i]:
            outs.append(th)
        th2, th1 = th1, th
    return torch.stack(outs, dim=2)
