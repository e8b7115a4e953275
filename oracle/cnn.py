"""CPU restatement of the learned closure network -- TEST INFRASTRUCTURE.

The reference has no learned closure: its force hook (src/opt/force.f90:28,
called from src/equations/ns_rho_rhou_p.f90:212) adds a user term to the
direct RHS.  This module restates the network that the library evaluates in
that place (include/cruna_hip.h, cruna_hip_cnn_*), forward and backward, in
numpy, so that the device's BF16 matrix-core kernels can be checked:

  features  x[p][c]  = bf16(f32(q[c][p] - m[c]) * f32(1 / in_std[c])), c < 5, zero to 16,
            m[c] the mean of field c over the grid
  layer l   z[p][co] = b[co] + sum_t sum_ci x[p + o_t][ci] bf16(W[t][co][ci])
            hidden:  x' = bf16(gelu(z)), z kept as bf16(z) for the backward pass
            last:    out = z (FP32)
  closure   force[v][p] = out[p][v] * s[v], s[v] = out_std[v], times
            f32(m[4] / p_ref) for v = 4 when p_ref > 0; then, for v < 4,
            minus the domain mean of force[v]
  loss      mean over p and the 5 real channels of (out - target / s)^2

with periodic 3x3x3 taps t = 9 (dz+1) + 3 (dy+1) + (dx+1), o_t = (dx, dy, dz)
along (i, j, k), and point p = i + S (j + S k).  With `bf16=False` every
rounding is skipped and the arithmetic is float64: that exact form is pinned
by finite differences and by a brute-force convolution (tests/test_oracle_cnn.py).

Only tests/ import this module.
"""
import numpy as np

CI = (16, 32, 32, 32)
CO = (32, 32, 32, 16)
REAL_IN = (5, 32, 32, 32)
REAL_OUT = (32, 32, 32, 5)
GELU_C, GELU_A = 0.7978845608028654, 0.044715


def layout():
    """(weight offset, bias offset) per layer in the flat parameter vector."""
    offs, o = [], 0
    for ci, co in zip(CI, CO):
        w = o
        o += 27 * co * ci
        offs.append((w, o))
        o += co
    return offs, o


def num_params():
    return layout()[1]


def unpack(params):
    """[(W (27, co, ci), b (co,))] per layer, float64."""
    offs, _ = layout()
    out = []
    for (w, b), ci, co in zip(offs, CI, CO):
        out.append((np.asarray(params[w:w + 27 * co * ci], dtype=np.float64).reshape(27, co, ci),
                    np.asarray(params[b:b + co], dtype=np.float64)))
    return out


def pack(layers):
    parts = []
    for W, b in layers:
        parts += [np.asarray(W, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64)]
    return np.concatenate(parts)


def to_bf16(x):
    """Round to bfloat16, nearest even (returned as float64)."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def _taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def _shift(a, dx, dy, dz):
    """b[k, j, i] = a[k + dz, j + dy, i + dx] (periodic); a is (S, S, S, C) indexed [k, j, i]."""
    return np.roll(a, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(x, W, b=None):
    """out[p][co] = b[co] + sum_t sum_ci x[p + o_t][ci] W[t][co][ci]."""
    S = x.shape[0]
    out = np.zeros((S, S, S, W.shape[1]))
    for t, dx, dy, dz in _taps():
        out += _shift(x, dx, dy, dz) @ W[t].T
    return out if b is None else out + b


def conv_transpose(d, W):
    """din[p][ci] = sum_t sum_co d[p - o_t][co] W[t][co][ci]."""
    S = d.shape[0]
    out = np.zeros((S, S, S, W.shape[2]))
    for t, dx, dy, dz in _taps():
        out += _shift(d, -dx, -dy, -dz) @ W[t]
    return out


def conv_wgrad(x, d):
    """dW[t][co][ci] = sum_p d[p][co] x[p + o_t][ci]."""
    C_in, C_out = x.shape[-1], d.shape[-1]
    dW = np.zeros((27, C_out, C_in))
    df = d.reshape(-1, C_out)
    for t, dx, dy, dz in _taps():
        dW[t] = df.T @ _shift(x, dx, dy, dz).reshape(-1, C_in)
    return dW


def _r(x, on):
    return to_bf16(x) if on else x


def _grid(f, S):
    """(S, S, S) field in Fortran order (i, j, k) -> [k, j, i] array."""
    return np.asarray(f).reshape((S, S, S), order="F").transpose(2, 1, 0)


def _ungrid(a):
    return np.asfortranarray(a.transpose(2, 1, 0))


def field_means(q):
    return np.array([q[..., c].mean() for c in range(5)])


def features(q, in_std, bf16=True):
    """Network input (S, S, S, 16) [k, j, i, c] of a state q (S, S, S, 5) Fortran order."""
    S = q.shape[0]
    m = field_means(q)
    x = np.zeros((S, S, S, 16))
    for c in range(5):
        d = _grid(q[..., c], S) - m[c]
        if bf16:
            v = d.astype(np.float32) * np.float32(1.0 / in_std[c])
        else:
            v = d / in_std[c]
        x[..., c] = _r(v, bf16)
    return x


def output_scales(q, out_std, p_ref, bf16=True):
    """s[v] of the closure force (FP32 products when bf16)."""
    ps = q[..., 4].mean() / p_ref if p_ref > 0 else 1.0
    if bf16:
        s = [np.float32(x) for x in out_std]
        s[4] = np.float32(out_std[4]) * np.float32(ps)
        return s
    s = list(out_std)
    s[4] = out_std[4] * ps
    return s


def forward(params, x, bf16=True):
    """(out (S,S,S,16), activations [x0..x3], pre-activations [z0..z2])."""
    layers = unpack(params)
    acts, zs = [x], []
    for l, (W, b) in enumerate(layers):
        z = conv(acts[-1], _r(W, bf16), b)
        if l == 3:
            return z, acts, zs
        if bf16:
            z = z.astype(np.float32).astype(np.float64)
        zs.append(_r(z, bf16))
        acts.append(_r(gelu(z), bf16))


def closure(params, q, in_std, out_std, p_ref=0.0, bf16=True):
    """force (S, S, S, 5) Fortran order of the state q."""
    out, _, _ = forward(params, features(q, in_std, bf16), bf16)
    s = output_scales(q, out_std, p_ref, bf16)
    S = q.shape[0]
    f = np.zeros((S, S, S, 5), order="F")
    for v in range(5):
        o = _ungrid(out[..., v])
        f[..., v] = (o.astype(np.float32) * s[v]).astype(np.float64) if bf16 else o * s[v]
        if v < 4:
            f[..., v] -= f[..., v].mean()
    return f


def loss_and_grads(params, x, target_n, bf16=True):
    """Loss and flat parameter gradient for features x and normalised targets
    target_n (S, S, S, 5) [k, j, i, v]."""
    layers = unpack(params)
    out, acts, zs = forward(params, x, bf16)
    S = x.shape[0]
    npts = S ** 3
    d = np.zeros_like(out)
    diff = out[..., :5] - target_n
    loss = float((diff ** 2).sum() / (npts * 5))
    d[..., :5] = _r(2.0 * diff / (npts * 5), bf16)
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, _ = layers[l]
        grads[l] = (conv_wgrad(acts[l], d), d.reshape(-1, d.shape[-1]).sum(axis=0))
        if l > 0:
            d = _r(conv_transpose(d, _r(W, bf16)) * gelu_grad(zs[l - 1]), bf16)
    return loss, pack(grads)


def normalised_targets(q, target, out_std, p_ref=0.0, bf16=True):
    """target / s [k, j, i, v] (the FP32 products of the device when bf16)."""
    S = target.shape[0]
    ps = q[..., 4].mean() / p_ref if p_ref > 0 else 1.0
    t = np.zeros((S, S, S, 5))
    for v in range(5):
        if bf16:
            inv = np.float32(1.0 / out_std[v])
            if v == 4:
                inv = inv / np.float32(ps)
            t[..., v] = _grid(target[..., v], S).astype(np.float32) * inv
        else:
            t[..., v] = _grid(target[..., v], S) / (out_std[v] * (ps if v == 4 else 1.0))
    return t
