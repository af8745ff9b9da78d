"""ORACLE — test infrastructure only: an FP64 CPU restatement of the router MLP's
forward pass, softmax cross-entropy gradient and Adam step (no counterpart in the
reference: kyto has no learned router). The device rounds each matrix product's
operands to bf16 and accumulates in f32; this restatement rounds the same operands
to bf16 and computes in f64, so the two agree to f32 accumulation error."""
import numpy as np


def bf16(a):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, d_in, hidden, n_out):
    sizes = [d_in * hidden, hidden, hidden * hidden, hidden, hidden * n_out, n_out]
    out, at = [], 0
    for s in sizes:
        out.append(np.asarray(params[at:at + s], np.float64))
        at += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(d_in, hidden), b1, W2.reshape(hidden, hidden), b2, W3.reshape(hidden, n_out), b3


def grad(params, x, t_idx, t_w, d_in, hidden, n_out):
    """Mean cross-entropy of softmax(MLP(x)) against sparse targets and its gradient
    with respect to every parameter (layout of `params`)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, d_in, hidden, n_out)
    B = x.shape[0]
    x = np.asarray(x, np.float64)
    z1 = bf16(x) @ bf16(W1) + b1
    h1 = np.maximum(z1, 0.0).astype(np.float32).astype(np.float64)
    z2 = bf16(h1) @ bf16(W2) + b2
    h2 = np.maximum(z2, 0.0).astype(np.float32).astype(np.float64)
    L = bf16(h2) @ bf16(W3) + b3
    m = L.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(L - m).sum(axis=1, keepdims=True))
    P = np.exp(L - lse)
    T = np.zeros_like(P)
    for r in range(B):
        for e in range(t_idx.shape[1]):
            T[r, t_idx[r, e]] += t_w[r, e]
    loss = float(-(T * (L - lse)).sum() / B)
    dL = ((P - T) / B).astype(np.float32).astype(np.float64)
    dW3 = bf16(h2).T @ bf16(dL)
    db3 = dL.sum(axis=0)
    dh2 = (bf16(dL) @ bf16(W3).T) * (h2 > 0)
    dh2 = dh2.astype(np.float32).astype(np.float64)
    dW2 = bf16(h1).T @ bf16(dh2)
    db2 = dh2.sum(axis=0)
    dh1 = (bf16(dh2) @ bf16(W2).T) * (h1 > 0)
    dh1 = dh1.astype(np.float32).astype(np.float64)
    dW1 = bf16(x).T @ bf16(dh1)
    db1 = dh1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return g, loss


def adam(w, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma and Ba 2015), bias-corrected, in f64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
