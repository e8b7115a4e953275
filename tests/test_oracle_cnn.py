"""Pins of the closure network's CPU restatement (oracle/cnn.py).

The network has no counterpart in the reference (its force hook,
src/opt/force.f90:28, is a user term), so the restatement is pinned by
what the mathematics fixes:
  * the convolution against a brute-force loop over points and taps;
  * the transposed convolution and the weight gradient as the adjoints of
    the convolution (<conv(x, W), y> = <x, conv_T(y, W)> = <W, wgrad(x, y)>);
  * the full backward pass against central finite differences of the loss
    (float64, no rounding);
  * BF16 rounding and GELU against torch's bfloat16 cast and
    gelu(approximate="tanh") with its autograd derivative.
"""
import numpy as np
import torch

from oracle import cnn


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    S, ci, co = 4, 3, 2
    x = rng.standard_normal((S, S, S, ci))
    W = rng.standard_normal((27, co, ci))
    b = rng.standard_normal(co)
    got = cnn.conv(x, W, b)
    ref = np.zeros_like(got)
    for k in range(S):
        for j in range(S):
            for i in range(S):
                for t in range(27):
                    dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
                    ref[k, j, i] += W[t] @ x[(k + dz) % S, (j + dy) % S, (i + dx) % S]
                ref[k, j, i] += b
    assert np.abs(got - ref).max() < 1e-12


def test_transpose_and_wgrad_are_adjoints():
    rng = np.random.default_rng(1)
    S, ci, co = 5, 4, 3
    x = rng.standard_normal((S, S, S, ci))
    y = rng.standard_normal((S, S, S, co))
    W = rng.standard_normal((27, co, ci))
    lhs = np.sum(cnn.conv(x, W) * y)
    assert abs(lhs - np.sum(x * cnn.conv_transpose(y, W))) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(W * cnn.conv_wgrad(x, y))) < 1e-10 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    S = 4
    params = rng.standard_normal(cnn.num_params()) * 0.05
    # padding entries stay zero as the library keeps them
    x = np.zeros((S, S, S, 16))
    x[..., :5] = rng.standard_normal((S, S, S, 5))
    tgt = rng.standard_normal((S, S, S, 5))
    loss, g = cnn.loss_and_grads(params, x, tgt, bf16=False)
    offs, n = cnn.layout()
    picks = []
    for l, (w, b) in enumerate(offs):
        for _ in range(6):
            t, o, i = rng.integers(27), rng.integers(cnn.REAL_OUT[l]), rng.integers(cnn.REAL_IN[l])
            picks.append(w + (t * cnn.CO[l] + o) * cnn.CI[l] + i)
        picks.append(b + rng.integers(cnn.REAL_OUT[l]))
    h = 1e-5
    for e in picks:
        pp, pm = params.copy(), params.copy()
        pp[e] += h
        pm[e] -= h
        fd = (cnn.loss_and_grads(pp, x, tgt, bf16=False)[0] - cnn.loss_and_grads(pm, x, tgt, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[e]) < 1e-6 * max(1.0, abs(g).max()), (e, fd, g[e])
    assert loss > 0


def test_bf16_and_gelu_match_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000), [0.0, -0.0, 1.0, 65504.0]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).float().numpy().astype(np.float64)
    assert np.array_equal(cnn.to_bf16(x), ref)
    z = torch.linspace(-8, 8, 4001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(z, approximate="tanh")
    y.sum().backward()
    zn = z.detach().numpy()
    assert np.abs(cnn.gelu(zn) - y.detach().numpy()).max() < 1e-14
    assert np.abs(cnn.gelu_grad(zn) - z.grad.numpy()).max() < 1e-13


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(4)
    p = rng.standard_normal(cnn.num_params())
    assert np.array_equal(cnn.pack(cnn.unpack(p)), p)
    assert cnn.num_params() == 27 * (16 * 32 + 32 * 32 * 2 + 32 * 16) + 32 * 3 + 16
