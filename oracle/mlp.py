"""FP64 numpy restatement of the learned-guess network -- TEST INFRASTRUCTURE ONLY.

The network of include/qcml.h: X -> SiLU(X W1 + b1) -> SiLU(H1 W2 + b2) -> H2 W3 + b3, mean squared
error loss, hand-written backward pass (no autodiff).  Parameters are one flat vector in the order
W1, b1, W2, b2, W3, b3, row-major.  ``bf16=True`` rounds every GEMM operand to BF16 (nearest even)
the way the GPU kernels stage them, so the GPU can be compared at FP32-accumulation precision.
No reference counterpart (the reference has no learned guess): the backward pass is pinned by
central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np


def bf16_round(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shapes(nin, width, nout):
    return [(nin, width), (width,), (width, width), (width,), (width, nout), (nout,)]


def unflatten(p, nin, width, nout):
    out, o = [], 0
    for s in shapes(nin, width, nout):
        k = int(np.prod(s))
        out.append(np.asarray(p[o:o + k], dtype=np.float64).reshape(s))
        o += k
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, dims, bf16=False):
    W1, b1, W2, b2, W3, b3 = unflatten(p, *dims)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Z1 = q(X) @ q(W1) + b1
    H1 = _silu(Z1)
    Z2 = q(H1) @ q(W2) + b2
    H2 = _silu(Z2)
    Y = q(H2) @ q(W3) + b3
    return Y, (Z1, H1, Z2, H2)


def loss_and_grad(p, X, T, dims, bf16=False):
    W1, b1, W2, b2, W3, b3 = unflatten(p, *dims)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (Z1, H1, Z2, H2) = forward(p, X, dims, bf16)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dY = 2.0 * (Y - T) / n
    gW3 = q(H2).T @ q(dY)
    gb3 = dY.sum(0)
    dZ2 = (q(dY) @ q(W3).T) * _dsilu(Z2)
    gW2 = q(H1).T @ q(dZ2)
    gb2 = dZ2.sum(0)
    dZ1 = (q(dZ2) @ q(W2).T) * _dsilu(Z1)
    gW1 = q(X).T @ q(dZ1)
    gb1 = dZ1.sum(0)
    return loss, np.concatenate([a.reshape(-1) for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
