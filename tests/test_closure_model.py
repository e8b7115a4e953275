"""Host logic of the closure workflow (no GPU): the fallback rule of
ClosureModel, the spectrum and box-average helpers of closure_study."""
import numpy as np

from public_amd import closure_study as CS


def _model(tmp_path, errors, margin=1.0):
    p = tmp_path / "m.npz"
    np.savez(p, S=64, factor=4, params=np.zeros(10, np.float32), in_std=np.ones(5), out_std=np.ones(5),
             p_ref=17.0, gain=[1, 1.2, 1.2, 1.2, 1], val_steps=np.arange(0, 301, 50), val_errors=np.array(errors))
    return CS.ClosureModel(str(p), threshold=0.10, margin=margin)


def test_fallback_rule_follows_validated_horizon(tmp_path):
    m = _model(tmp_path, [0.0, 0.01, 0.03, 0.08, 0.12, 0.2, 0.3])
    assert m.usable(50) and m.usable(150) and m.usable(120)
    assert not m.usable(151) and not m.usable(200) and not m.usable(300)
    assert not m.usable(400)  # beyond the validated horizon
    assert m.validated_error(120) == 0.08
    assert m.horizon() == 150
    assert np.allclose(m.gain, [1, 1.2, 1.2, 1.2, 1])


def test_fallback_rule_margin(tmp_path):
    # the default margin of 1/2: the closure only while the validated error
    # is within half the threshold
    m = _model(tmp_path, [0.0, 0.01, 0.03, 0.08, 0.12, 0.2, 0.3], margin=0.5)
    assert m.usable(50) and m.usable(100) and not m.usable(101) and not m.usable(150)
    assert m.horizon() == 100
    # a curve that dips back under the bar after leaving it: the horizon ends
    # at the first checkpoint beyond the bar
    m = _model(tmp_path, [0.0, 0.01, 0.06, 0.04, 0.03, 0.2, 0.3], margin=0.5)
    assert m.horizon() == 50
    assert m.usable(50) and not m.usable(100) and not m.usable(200)


def test_velocity_spectrum_of_a_single_mode():
    n = 16
    x = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0] = 1.0
    q[..., 1] = np.cos(3 * Y)  # u = cos(3y): energy 1/4 in shell 3
    E = CS.velocity_spectrum(q)
    assert abs(E[3] - 0.25) < 1e-12 and abs(E.sum() - 0.25) < 1e-12
    assert CS.spectrum_error(q, q) == 0.0


def test_box_average():
    rng = np.random.default_rng(0)
    q = np.asfortranarray(rng.standard_normal((8, 8, 8, 5)))
    b = CS.box_average(q, 4)
    assert b.shape == (2, 2, 2, 5)
    assert np.isclose(b[1, 0, 1, 3], q[4:8, 0:4, 4:8, 3].mean())
