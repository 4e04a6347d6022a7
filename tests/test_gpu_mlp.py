"""GPU: the learned-guess network's own kernels (libqcml.so) against the FP64 numpy oracle.

* GEMM: every transpose combination and ragged sizes, against numpy on the BF16-rounded operands
  (the kernel rounds FP32 operands to BF16 on load and accumulates in FP32): 1e-5 relative.
* Gradients of the MSE loss for every parameter against the oracle's hand-written backward pass
  with the same BF16 rounding of GEMM operands: 2e-3 relative (norm-wise, FP32 accumulation and
  activation-rounding differences), cosine > 0.99999.  The oracle's backward pass itself is pinned
  by finite differences in test_mlp_oracle.py.
* Training: Adam steps drive the loss of a smooth regression task down by 10x.
"""
import numpy as np
import pytest

import oracle.mlp as om

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (37, 70, 45), (130, 17, 300), (1, 256, 256)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm(require_gpu, M, N, K, ta, tb):
    import torch
    from qchem_amd import mlp
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0))
    mlp.gemm(dA, dB, dC, transA=ta, transB=tb, alpha=0.5, beta=2.0)
    Ar, Br = om.bf16_round(A), om.bf16_round(B)
    ref = 0.5 * ((Ar.T if ta else Ar) @ (Br.T if tb else Br)) + 2.0 * C0
    got = dC.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-5 * max(1.0, np.abs(ref).max()) * np.sqrt(K)


def test_gradients_vs_oracle(require_gpu):
    import torch
    from qchem_amd import mlp
    dims = (45, 256, 333)
    net = mlp.Net(*dims, max_batch=64, seed=11)
    rng = np.random.default_rng(12)
    X = rng.normal(size=(48, dims[0])).astype(np.float32)
    T = rng.normal(size=(48, dims[2])).astype(np.float32)
    loss, g = net.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    p = net.get_params().astype(np.float64)
    lref, gref = om.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), dims, bf16=True)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    print("loss %.6f ref %.6f grad rel %.2e cos %.8f" % (loss, lref, rel, cos))
    assert abs(loss - lref) < 1e-5 * lref
    assert rel < 2e-3 and cos > 0.99999
    # per-layer: every block of the gradient, not just the largest, agrees
    o = 0
    for s in om.shapes(*dims):
        k = int(np.prod(s))
        assert np.linalg.norm(g[o:o + k] - gref[o:o + k]) < 5e-3 * np.linalg.norm(gref[o:o + k]) + 1e-9
        o += k


def test_training_reduces_loss(require_gpu):
    import torch
    from qchem_amd import mlp
    rng = np.random.default_rng(13)
    X = rng.uniform(-1, 1, size=(256, 8)).astype(np.float32)
    Wt = rng.normal(size=(8, 32))
    T = np.tanh(X @ Wt).astype(np.float32)
    dX, dT = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    net = mlp.Net(8, 64, 32, max_batch=256, seed=5)
    l0 = net.train_step(dX, dT, 1e-3, want_loss=True)
    for _ in range(300):
        net.train_step(dX, dT, 1e-3)
    l1 = net.train_step(dX, dT, 1e-3, want_loss=True)
    print("loss %.5f -> %.5f" % (l0, l1))
    assert l1 < 0.1 * l0
