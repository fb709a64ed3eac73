"""The CPU reference of the guess network pinned independently of the
device: the tap convention by a delta response, the backward pass by
central finite differences in FP64, and the training-field generator."""
import numpy as np
import pytest
import torch

import nn_reference as R


def test_tap_convention_delta_response():
    """W[t][0][0] = 1 for the single tap t = (dx, dy, dz) moves a unit impulse
    at cell c to cell c - (dx, dy, dz): the layer reads input (i+dx, j+dy, k+dz)."""
    for dx, dy, dz in [(1, 0, 0), (0, -1, 0), (0, 0, 1), (-1, 1, -1)]:
        t = 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1)
        w = np.zeros((27, 1, 1), dtype=np.float32)
        w[t, 0, 0] = 1.0
        x = torch.zeros(1, 1, 6, 6, 6)
        x[0, 0, 3, 3, 3] = 1.0
        y = torch.nn.functional.conv3d(x, R.torch_weight(w), padding=1)[0, 0]
        idx = np.argwhere(y.numpy() == 1.0)
        assert idx.tolist() == [[3 - dx, 3 - dy, 3 - dz]]


def test_reference_gradients_by_finite_differences():
    n = 16
    rng = np.random.default_rng(0)
    lp = []
    for cin, cout, _ in R.LAYERS:
        lp.append((rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin)),
                   0.1 * rng.standard_normal(cout)))
    feat = rng.standard_normal((n, n, n, 2))
    mono = rng.standard_normal((n, n, n))
    target = rng.standard_normal((n, n, n))
    loss, _, grads = R.loss_and_grads(lp, feat, mono, target, rounding=False,
                                      dtype=torch.float64)
    for l, ci, kind in [(0, (0, 3, 13), "w"), (5, (7, 2, 4), "w"), (11, (5, 0, 26), "w"),
                        (8, 3, "b"), (11, 0, "b")]:
        w, b = lp[l]
        h = 1e-6
        if kind == "w":
            c, o, t = ci
            wp, wm = w.copy(), w.copy()
            wp[t, c, o] += h
            wm[t, c, o] -= h
            lp_p = list(lp); lp_p[l] = (wp, b)
            lp_m = list(lp); lp_m[l] = (wm, b)
            g = grads[l][0][t, c, o]
        else:
            bp, bm = b.copy(), b.copy()
            bp[ci] += h
            bm[ci] -= h
            lp_p = list(lp); lp_p[l] = (w, bp)
            lp_m = list(lp); lp_m[l] = (w, bm)
            g = grads[l][1][ci]
        fp = R.loss_and_grads(lp_p, feat, mono, target, rounding=False, dtype=torch.float64)[0]
        fm = R.loss_and_grads(lp_m, feat, mono, target, rounding=False, dtype=torch.float64)[0]
        fd = (fp - fm) / (2 * h)
        assert g == pytest.approx(fd, rel=1e-5, abs=1e-9), (l, kind, ci)


def test_mono_field_closed_form():
    m = R.mono_field(8, np.array([1.0, 0.5, 0.5, 0.5]))
    # cells next to the centre: r = h sqrt(3) / 2, eps = 1.5 h, h = 1/8
    h = 1.0 / 8
    expect = -1.0 / np.sqrt(0.75 * h * h + 2.25 * h * h)
    assert m[3, 3, 3] == pytest.approx(expect, rel=1e-6)
    assert m[4, 4, 4] == pytest.approx(expect, rel=1e-6)
    assert np.allclose(m, m[::-1, ::-1, ::-1])


def test_training_fields_are_seeded():
    from hydrocode_astro1_amd.nn import field_params
    a = field_params(7)
    b = field_params(7)
    c = field_params(8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[0], c[0]) or not np.array_equal(a[1], c[1])
    blobs, modes, floor_rho, sigma = a
    assert 1 <= len(blobs) <= 8
    assert np.all((blobs[:, :3] >= 0.2) & (blobs[:, :3] <= 0.8))
    assert np.all((blobs[:, 3] >= 0.1) & (blobs[:, 3] <= 10.0))
    assert np.all((blobs[:, 4] >= 0.03) & (blobs[:, 4] <= 0.15))
    # lognormal factor: std of delta equals sigma
    assert np.sqrt(0.5 * np.sum(modes[:, 4] ** 2)) == pytest.approx(sigma, rel=1e-12)


def test_gradient_rounding_is_identity_forward_bf16_backward():
    """gbf16 leaves values alone and rounds the incoming gradient to the
    nearest BF16 (ties to even): 1 + 2^-8 is a tie that rounds down to 1,
    1 + 3 * 2^-8 rounds up to 1 + 2^-6."""
    x = torch.tensor([0.5, -2.0, 3.0], requires_grad=True)
    y = R.gbf16(x)
    assert torch.equal(y, x)
    c = torch.tensor([1.0 + 2.0 ** -8, -(1.0 + 3 * 2.0 ** -8), 0.1])
    (y * c).sum().backward()
    expect = torch.tensor([1.0, -(1.0 + 2.0 ** -6), 0.10009765625])
    assert torch.equal(x.grad, expect)
    # disabled: the gradient passes unchanged
    x.grad = None
    (R.gbf16(x, enabled=False) * c).sum().backward()
    assert torch.equal(x.grad, c)


def test_correction_pair_helpers():
    """The start correction's host helpers: the 7-point Laplacian is exact on
    a quadratic (6 on the interior, 0 on the boundary), the f^3 average keeps
    the mean, and C is the median ratio rms(e) / (rms(r) L^2)."""
    from hydrocode_astro1_amd import nn
    n = 16
    h = 1.0 / n
    X = (np.arange(n) + 0.5) * h
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    lap = nn.laplacian(x * x + y * y + z * z, h)
    assert np.allclose(lap[1:-1, 1:-1, 1:-1], 6.0, rtol=1e-9)
    assert np.all(lap[0] == 0) and np.all(lap[:, :, -1] == 0)
    a = np.random.default_rng(1).standard_normal((n, n, n))
    avg = nn.average(a, 4)
    assert avg.shape == (4, 4, 4)
    assert avg.mean() == pytest.approx(a.mean(), abs=1e-12)
    assert avg[1, 2, 3] == pytest.approx(a[4:8, 8:12, 12:16].mean())
    e = nn.zero_boundary(a.copy())
    assert np.all(e[:, 0] == 0) and np.all(e[-1] == 0) and np.all(e[1:-1, 1:-1, 1:-1] == a[1:-1, 1:-1, 1:-1])
    R_ = [a, 2 * a, 0 * a, a]
    E_ = [3 * a, 6.5 * a, a, 2 * a]
    assert nn.correction_scale(R_, E_) == pytest.approx(3.0)
    assert nn.correction_scale(R_, E_, box_len=2.0) == pytest.approx(0.75)
