"""CPU oracle of the warm-start multilayer perceptron (test infrastructure only: used by
tests/ as the checker of the HIP kernels in multiphaselb_amd/csrc/mlp.hip, never by the
product).  The network has no counterpart in the reference; this restates the network
the kernels compute:

  level 0   x padded with zeros to 32 columns, rounded to BF16
  layer l   z = a_l W_l + b_l; hidden layers a_{l+1} = tanh(z) kept in FP32 and
            rounded to BF16 as the next product's operand; the last layer is linear
  loss      L = sum (y - t)^2 / (B n_out)
  backward  dz_L = 2 (y - t) / (B n_out); dW_l = a_l^T dz_{l+1}; db_l = sum_b dz_{l+1};
            dz_l = (dz_{l+1} W_l^T) * (1 - a_l^2), every product operand in BF16

Products are summed in float64 here (the GPU accumulates in FP32), so the two agree to
FP32 rounding wherever both see the same BF16 operands.  With bf16=False every
rounding is dropped and the arithmetic is float64 throughout: that mode is pinned
against central finite differences of the loss.
"""
import numpy as np

IN_PAD = 32


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def layers(params, widths):
    """Split the flat parameter vector (per layer W [in][out] then b [out])."""
    out, o = [], 0
    for l in range(len(widths) - 1):
        n_in, n_out = widths[l], widths[l + 1]
        W = params[o:o + n_in * n_out].reshape(n_in, n_out)
        o += n_in * n_out
        b = params[o:o + n_out]
        o += n_out
        out.append((W, b))
    return out


def forward(params, widths, x, bf16_ops=True):
    """Returns the per-level operands, FP32 activations and the output."""
    rnd = bf16 if bf16_ops else (lambda v: np.asarray(v, np.float64))
    fl = (lambda v: np.asarray(v, np.float32)) if bf16_ops else (lambda v: np.asarray(v, np.float64))
    L = len(widths) - 1
    B = x.shape[0]
    a0 = np.zeros((B, IN_PAD), np.float64)
    a0[:, :widths[0]] = x
    ops = [rnd(a0).astype(np.float64)]
    acts = [None]
    y = None
    for l, (W, b) in enumerate(layers(np.asarray(params, np.float64), widths)):
        Wp = np.zeros((ops[l].shape[1], W.shape[1]))
        Wp[:W.shape[0]] = W
        z = ops[l] @ rnd(Wp).astype(np.float64) + np.asarray(fl(b), np.float64)
        if l < L - 1:
            a = fl(np.tanh(z))
            acts.append(np.asarray(a, np.float64))
            ops.append(rnd(a).astype(np.float64))
        else:
            y = np.asarray(fl(z), np.float64)
    return ops, acts, y


def loss_grad(params, widths, x, t, bf16_ops=True):
    """Loss and flat gradient (same layout as the parameters)."""
    rnd = bf16 if bf16_ops else (lambda v: np.asarray(v, np.float64))
    fl = (lambda v: np.asarray(v, np.float32)) if bf16_ops else (lambda v: np.asarray(v, np.float64))
    ops, acts, y = forward(params, widths, x, bf16_ops)
    B, nout = y.shape
    L = len(widths) - 1
    d = y - t
    loss = float(np.sum(d * d) / (B * nout))
    dz = np.asarray(fl(2.0 * d / (B * nout)), np.float64)
    lay = layers(np.asarray(params, np.float64), widths)
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        W, b = lay[l]
        dzb = rnd(dz).astype(np.float64)
        dW = ops[l].T @ dzb
        db = np.asarray(fl(dz.sum(axis=0)), np.float64)
        grads[l] = (dW[:W.shape[0]], db)
        if l > 0:
            Wp = rnd(W).astype(np.float64)
            dA = dzb @ Wp.T
            dz = np.asarray(fl(dA * (1.0 - acts[l] * acts[l])), np.float64)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return loss, flat


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update with bias correction at step t (1-based), float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
