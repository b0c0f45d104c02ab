"""numpy restatement of the learned-axis MLP (forward, loss, backward).

TEST INFRASTRUCTURE ONLY (tests/ use it to check the HIP kernels' outputs and
gradients).  Same parameter layout and semantics as gsap_amd/csrc/k_mlp.hip:
bf16 rounding (round to nearest even) of inputs, weights and hidden
activations, accumulation in higher precision, loss = mean over rows of the
expected regret sum_c softmax_c(kappa (u.a_c)^2) r_c over the 64
candidate axes, plus lambda (y3 - t64)^2; gradients with respect to the
bf16-rounded weights.
Pinned by finite differences on the unrounded (bf16=False) network.
"""
import numpy as np

F, H, O = 64, 128, 4
OFF_W1, OFF_B1 = 0, F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * O
N_PARAMS = OFF_B3 + O


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(H, O), p[OFF_B3:OFF_B3 + O])


def forward(p, x, use_bf16=True):
    r = (lambda a: bf16(a).astype(np.float64)) if use_bf16 else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    xb = r(x)
    h1 = r(np.maximum(xb @ W1 + b1, 0.0))
    h2 = r(np.maximum(h1 @ W2 + b2, 0.0))
    y = h2 @ W3 + b3
    return y, (xb, h1, h2, W1, W2, W3)


def loss_grad(p, x, t, cand, lam=0.05, kappa=30.0, use_bf16=True):
    """t: rows of 64 candidate regrets then log(1 + pairs); cand: 64 x 3 axes."""
    y, (xb, h1, h2, W1, W2, W3) = forward(p, x, use_bf16)
    t = np.asarray(t, np.float64)
    A = np.asarray(cand, np.float64)
    B = len(x)
    u = y[:, :3]
    nu = np.sqrt((u * u).sum(1)) + 1e-6
    uh = u / nu[:, None]
    dc = uh @ A.T                                   # B x 64
    s = kappa * dc * dc
    e = np.exp(s - s.max(1, keepdims=True))
    pc = e / e.sum(1, keepdims=True)
    r = t[:, :64]
    la = (pc * r).sum(1)
    ec = y[:, 3] - t[:, 64]
    loss = float(np.mean(la + lam * ec * ec))
    gs = kappa * pc * (r - la[:, None]) * 2 * dc    # B x 64
    g = gs @ A                                      # dL/du_hat
    gr = (g * uh).sum(1)
    dy = np.zeros_like(y)
    dy[:, :3] = (g - gr[:, None] * uh) / nu[:, None] / B
    dy[:, 3] = 2 * lam * ec / B
    dh2 = (dy @ W3.T) * (h2 > 0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    grads = np.concatenate([(xb.T @ dh1).ravel(), dh1.sum(0), (h1.T @ dh2).ravel(), dh2.sum(0),
                            (h2.T @ dy).ravel(), dy.sum(0)])
    return loss, grads, y
