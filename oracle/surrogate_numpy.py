"""NumPy restatement of the convolutional coarse-propagator surrogate.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The surrogate has no counterpart in the reference (whose learned operators
are Keras models, pararealml/operators/ml/supervised/supervised_ml_operator.py
:91-128 defines only the autoregressive protocol it follows).  This module
restates the device network so that its forward pass and every gradient can
be checked:

* six 3x3 "same" convolutions with zero padding, NHWC layout:
  C_in -> 64 -> 64 -> 64 -> 64 -> 64 -> C_out, tanh-form GELU after the
  first five, no activation after the last;
* the input is normalised by the per-sample scale s = max |u| and the
  prediction is ``u + s * net(x)`` (an increment step; the identity part is
  not rounded);
* the loss is the mean over pixels and channels of ``(net(x) - d)^2`` with
  the normalised increment ``d = (t - u) / s``.

Arithmetic is FP64, with BF16 round-to-nearest-even applied exactly where the
device stores BF16 (weights of layers 1-5, activations, pre-activations,
activation gradients; the first layer reads the normalised input and its
weights unrounded, as the device does in FP32), so device and oracle differ
only by FP32-vs-FP64 accumulation.

Pinning: ``conv3x3`` is checked against a direct sum on tiny inputs and
against linearity; the backward passes against finite differences of the
FP64 (unrounded) forward; ``gelu`` against its closed form at known points
(tests/test_oracle_surrogate.py); ``adam_step`` against the closed form of
the first bias-corrected step, lr g / (|g| + eps), and of every step under a
constant gradient (tests/test_oracle_pins.py).
"""

import numpy as np

HIDDEN = 64
N_LAYERS = 6
GELU_C = np.sqrt(2.0 / np.pi)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    bits = a.view(np.uint32).astype(np.uint64)
    rounding = ((bits >> 16) & 1) + 0x7FFF
    out = ((bits + rounding) >> 16) << 16
    out = out.astype(np.uint32)
    res = out.view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    if np.any(nan):
        res[nan] = np.nan
    return res


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z**3)))


def gelu_grad(z):
    u = GELU_C * (z + 0.044715 * z**3)
    t = np.tanh(u)
    du = GELU_C * (1.0 + 3 * 0.044715 * z**2)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * du


def conv3x3(x: np.ndarray, w: np.ndarray, b=None) -> np.ndarray:
    """x: (B, H, W, Ci); w: (3, 3, Ci, Co) indexed [kh, kw, ci, co] with the
    output at (h, w) reading input (h + kh - 1, w + kw - 1), zero padded."""
    bsz, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((bsz, h, wd, w.shape[-1]))
    for kh in range(3):
        for kw in range(3):
            out += np.einsum("bhwc,cd->bhwd", xp[:, kh : kh + h, kw : kw + wd, :], w[kh, kw])
    if b is not None:
        out += b
    return out


def conv3x3_input_grad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of conv3x3 with respect to its input."""
    bsz, h, wd, _ = dy.shape
    dyp = np.pad(dy, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((bsz, h, wd, w.shape[2]))
    for kh in range(3):
        for kw in range(3):
            # input (h', w') feeds output (h' - kh + 1, w' - kw + 1)
            sl = dyp[:, 2 - kh : 2 - kh + h, 2 - kw : 2 - kw + wd, :]
            out += np.einsum("bhwd,cd->bhwc", sl, w[kh, kw])
    return out


def conv3x3_weight_grad(x: np.ndarray, dy: np.ndarray) -> np.ndarray:
    bsz, h, wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((3, 3, ci, dy.shape[-1]))
    for kh in range(3):
        for kw in range(3):
            out[kh, kw] = np.einsum("bhwc,bhwd->cd", xp[:, kh : kh + h, kw : kw + wd, :], dy)
    return out


def init_params(c_in: int, c_out: int, seed: int = 0, hidden: int = HIDDEN):
    """He-style initial weights (BF16-representable) and zero biases; the
    last layer starts at zero so the untrained surrogate is the identity."""
    rng = np.random.RandomState(seed)
    chans = [c_in] + [hidden] * (N_LAYERS - 1) + [c_out]
    params = []
    for layer in range(N_LAYERS):
        ci, co = chans[layer], chans[layer + 1]
        std = np.sqrt(2.0 / (9 * ci))
        w = rng.normal(0.0, std, (3, 3, ci, co))
        if layer == N_LAYERS - 1:
            w *= 0.0
        params.append((bf16(w), np.zeros(co)))
    return params


def forward(params, x, rounding=True):
    """Returns (y, cache) for normalised input x (B, H, W, C)."""
    r = bf16 if rounding else (lambda v: v)
    a = x  # the first layer reads the normalised input unrounded
    cache = [a]
    zs = []
    for layer, (w, b) in enumerate(params):
        z = conv3x3(a, w, b)
        if layer < N_LAYERS - 1:
            zr = r(z)
            zs.append(zr)
            a = r(gelu(z))
            cache.append(a)
        else:
            zs.append(z)
            y = z
    return y, (cache, zs)


def loss_and_grads(params, x, inc, rounding=True):
    """MSE loss of ``net(x)`` against the normalised increment ``inc`` and
    the parameter gradients, following the device's rounding points."""
    r = bf16 if rounding else (lambda v: v)
    y, (acts, zs) = forward(params, x, rounding)
    n = y.size
    diff = y - inc
    loss = float(np.mean(diff**2))
    dz = r(2.0 * diff / n)
    grads = [None] * N_LAYERS
    for layer in range(N_LAYERS - 1, -1, -1):
        w, _ = params[layer]
        a_in = acts[layer]
        dw = conv3x3_weight_grad(a_in, dz)
        db = dz.sum(axis=(0, 1, 2))
        grads[layer] = (dw, db)
        if layer > 0:
            da = conv3x3_input_grad(dz, w)
            dz = r(da * gelu_grad(zs[layer - 1]))
    return loss, grads


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on FP64 masters; returns new (BF16-rounded weights,
    biases) and the updated state (masters, m, v, t)."""
    masters, m, v, t = state
    t = t + 1
    new_params, new_masters, new_m, new_v = [], [], [], []
    for layer, ((mw, mb), (gw, gb), (mmw, mmb), (vw, vb)) in enumerate(zip(masters, grads, m, v)):
        out = []
        for p, g, mm, vv in ((mw, gw, mmw, vw), (mb, gb, mmb, vb)):
            mm = beta1 * mm + (1 - beta1) * g
            vv = beta2 * vv + (1 - beta2) * g * g
            mhat = mm / (1 - beta1**t)
            vhat = vv / (1 - beta2**t)
            p = p - lr * mhat / (np.sqrt(vhat) + eps)
            out.append((p, mm, vv))
        (pw, mw2, vw2), (pb, mb2, vb2) = out
        new_masters.append((pw, pb))
        new_m.append((mw2, mb2))
        new_v.append((vw2, vb2))
        # the first layer computes with its FP32 master weights
        new_params.append((pw if layer == 0 else bf16(pw), pb))
    return new_params, (new_masters, new_m, new_v, t)


def predict(params, u):
    """One surrogate step on raw states u (B, H, W, C): u + s net(u / s)."""
    s = np.maximum(np.abs(u).reshape(u.shape[0], -1).max(axis=1), 1e-30)
    x = u / s[:, None, None, None]
    y, _ = forward(params, x)
    return u + s[:, None, None, None] * y
