"""NumPy float64 restatement of the proposal network of the learned sampler
(test infrastructure only): forward, masked softmax cross-entropy, backward
and one Adam step, on the same weights the GPU uses (bf16-rounded copies for
the products, fp32 masters for the optimiser).  Used to check the HIP
kernels' gradients.

    h = relu(x W1^T + b1),  z = h W2^T + b2,
    loss_i = logsumexp(z_i[lo_i:]) - z_i[t_i]
"""
import numpy as np


def bf16_round(w: np.ndarray) -> np.ndarray:
    """Round fp32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(seq: np.ndarray, m: int) -> tuple:
    """Teacher-forcing inputs of every (sequence, step) pair: row b*n + p."""
    B, n = seq.shape
    X = np.zeros((B * n, m + n))
    lo = np.zeros(B * n, dtype=np.int64)
    tgt = np.zeros(B * n, dtype=np.int64)
    for b in range(B):
        for p in range(n):
            r = b * n + p
            for q in range(p):
                X[r, seq[b, q]] += 1
            X[r, m + p] = 1
            lo[r] = seq[b, p - 1] if p else 0
            tgt[r] = seq[b, p]
    return X, lo, tgt


def loss_and_grads(W1, b1, W2, b2, X, lo, tgt, scale=1.0):
    W1b, W2b = bf16_round(W1), bf16_round(W2)
    Hpre = X @ W1b.T + b1
    H = np.maximum(Hpre, 0)
    Z = H @ W2b.T + b2
    R, m = Z.shape
    loss = np.zeros(R)
    dZ = np.zeros_like(Z)
    for i in range(R):
        z = Z[i, lo[i]:]
        mx = z.max()
        lse = mx + np.log(np.exp(z - mx).sum())
        loss[i] = lse - Z[i, tgt[i]]
        q = np.zeros(m)
        q[lo[i]:] = np.exp(Z[i, lo[i]:] - lse)
        q[tgt[i]] -= 1
        dZ[i] = q * scale
    dW2 = dZ.T @ H
    db2 = dZ.sum(axis=0)
    dH = dZ @ W2b
    dH[Hpre <= 0] = 0
    dW1 = dH.T @ X
    db1 = dH.sum(axis=0)
    return loss, (dW1, db1, dW2, db2)


def adam(w, g, m1, m2, lr, b1, b2, eps, step):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    w = w - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return w, m1, m2
