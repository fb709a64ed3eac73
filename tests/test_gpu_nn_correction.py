"""The start-correction network (kind 1 of the learned-guess network) on the
device: its features and forward pass against the torch CPU reference
(tests/nn_reference.py), the corrected start that POISSON_NN_FP32_DC builds
from it against the same start built in numpy, the weights file, and the
committed weights on a held-out explosion run."""
import os

import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics, nn
import nn_reference as R

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def to_bf16(x):
    """Round-to-nearest-even FP32 -> BF16, widened back to FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def smooth_residual(n, seed):
    rng = np.random.default_rng(seed)
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    r = (np.sin(2 * np.pi * x + rng.uniform(0, 6)) * np.cos(3 * np.pi * y) * np.sin(np.pi * z)
         + 0.3 * rng.standard_normal((n, n, n)))
    return (1e3 * r).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1])
def test_correction_features_and_forward_match_torch(seed):
    n = 32
    m = nn.Model(seed=seed)
    m.set_kind(nn.KIND_CORRECTION, 0.02)
    p = m.get_params()
    rng = np.random.default_rng(200 + seed)
    p += (0.02 * rng.standard_normal(p.size)).astype(np.float32) * (p != 0)
    m.set_params(p)
    r = smooth_residual(n, seed)
    target = (0.1 * rng.standard_normal((n, n, n))).astype(np.float32)
    loss, out, feat, grads = m.forward_backward(r, target, np.zeros(4))
    # features: r / s and -r / s in BF16, s the root mean square of r
    s = np.sqrt(np.mean(np.square(r.astype(np.float64))))
    a = (r * np.float32(1.0 / s)).astype(np.float32)
    assert np.array_equal(feat[..., 0], to_bf16(a))
    assert np.array_equal(feat[..., 1], -feat[..., 0])
    # no monopole term in this kind
    rloss, rout, rgrads = R.loss_and_grads(m.layer_params(p), feat, np.zeros((n, n, n), np.float32),
                                           target, grad_rounding=True)
    assert rel_l2(out, rout) < 1e-2
    assert loss == pytest.approx(rloss, rel=3e-3)
    # the gradients differ where FP32 sums taken in another order land on the
    # other side of a BF16 rounding boundary; the bias gradients of the 4^3
    # level sum only 64 voxels, so one such flip moves them most (measured:
    # 3.0e-2 relative L2 for the bias gradient of layer 6 with seed 1)
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(m.layer_params(grads), rgrads)):
        assert rel_l2(gw, rw) < 3e-2, f"layer {l} weight gradient"
        assert rel_l2(gb, rb) < 6e-2, f"layer {l} bias gradient"
        cos = float(np.dot(gw.ravel(), rw.ravel()) /
                    (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        assert cos > 0.9995, f"layer {l} weight gradient direction"


def test_kinds_in_the_weights_file_and_on_the_solver(tmp_path):
    a = nn.Model(seed=5)
    a.set_kind(nn.KIND_CORRECTION, 3.25e-3)
    f = tmp_path / "corr.bin"
    a.save(str(f))
    b = nn.Model(seed=6, path=str(f))
    assert b.kind() == (nn.KIND_CORRECTION, 3.25e-3)
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    with open(f, "rb") as fh:
        assert fh.read(8) == b"HYDNN003"
    g0 = tmp_path / "guess.bin"
    nn.Model(seed=7).save(str(g0))
    with open(g0, "rb") as fh:
        assert fh.read(8) == b"HYDNN002"
    b.load(str(g0))  # a potential-guess file sets kind 0 again
    assert b.kind()[0] == nn.KIND_GUESS
    with pytest.raises(ValueError):
        b.set_kind(nn.KIND_CORRECTION, -1.0)
    g = H.Gravity(32, 32, 32, G=1.0)
    with pytest.raises(ValueError):
        g.set_nn_correction(b)  # kind 0
    with pytest.raises(ValueError):
        g.set_nn_model(a)  # kind 1
    g.set_nn_correction(a)
    g.set_nn_model(b)
    g.set_nn_correction(None)


def interior_residual(phi, f, h):
    r = np.zeros_like(phi)
    r[1:-1, 1:-1, 1:-1] = f[1:-1, 1:-1, 1:-1] - nn.laplacian(phi, h)[1:-1, 1:-1, 1:-1]
    return r


def cubic_weights(n, nc):
    """W[i][c]: the interpolation from nc cell centres to n = f nc cell
    centres by cubic convolution (Keys, a = -1/2) at coarse coordinate
    x = (i + 1/2) / f - 1/2, the coarse values continued oddly past the edges
    (value -e[c] at cell -1 - c and at 2 nc - 1 - c)."""
    f = n // nc
    W = np.zeros((n, nc))
    for i in range(n):
        x = (i + 0.5) / f - 0.5
        a = int(np.floor(x))
        for c in range(a - 1, a + 3):
            d = abs(x - c)
            w = 1.5 * d ** 3 - 2.5 * d ** 2 + 1 if d < 1 else -0.5 * d ** 3 + 2.5 * d ** 2 - 4 * d + 2
            if c < 0:
                W[i, -c - 1] -= w
            elif c >= nc:
                W[i, 2 * nc - c - 1] -= w
            else:
                W[i, c] += w
    return W


def interpolate(W, e):
    t = np.einsum("kc,abc->abk", W, e)
    t = np.einsum("jb,abk->ajk", W, t)
    return np.einsum("ia,ajk->ijk", W, t)


@pytest.mark.parametrize("net", ["constant 0.7", "constant -0.4", "random 3"])
def test_corrected_start_matches_numpy(net):
    """The corrected start of POISSON_NN_FP32_DC is the warm start plus
    C L^2 s P(out) on the interior cells: out the network's output on the
    warm residual averaged 2x (s its root mean square), P cubic convolution
    from 64^3 to 128^3.  Its residual, as the solver reports it, matches the
    same start built in numpy; the solve keeps the smaller residual and
    converges.  With every weight zero and the last bias beta the output is
    beta everywhere; a random network checks the interpolation's taps."""
    n, nc = 128, 64
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = 1e-3 + np.exp(-((x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.55) ** 2) / 0.01)

    def solver(tol):
        g = H.Gravity(n, n, n, G=1.0)
        g.set_lmax(4)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        g.set_params(1.5, 400, tol)
        g.set_mode(H.POISSON_FP32_DC)
        g.set_dc_inner(H.DC_INNER_MULTIGRID)
        return g

    ref = solver(1e-11)
    assert ref.solve_potential(rho) == 0
    phi_ref = ref.phi()
    # warm start: the solution plus a smooth error, same boundary values
    delta = nn.zero_boundary(0.01 * np.abs(phi_ref).max() * np.sin(np.pi * x) ** 2
                             * np.sin(2 * np.pi * y) * np.sin(np.pi * z))
    phi_w = phi_ref + delta
    h = 1.0 / n
    f = 4.0 * np.pi * rho
    rw = interior_residual(phi_w, f, h)
    ravg = nn.average(rw.astype(np.float32).astype(np.float64), n // nc)

    kind, arg = net.split()
    m = nn.Model(seed=0)
    cs = 2.5e-3
    if kind == "constant":
        p = np.zeros(m.nparam, dtype=np.float32)
        p[H.lib().nn_model_param_offset(m.h, 11, 1)] = float(arg)
        # the solver applies the correction in FP32 (the inner solve's
        # unknown) and its Laplacian in FP32
        tol = 3e-5
    else:
        p = m.get_params()
        rng = np.random.default_rng(int(arg))
        p += (0.05 * rng.standard_normal(p.size)).astype(np.float32) * (p != 0)
        # the device averages in FP32 in its own order: a feature may round
        # to the neighbouring BF16 value, which moves the output slightly
        tol = 2e-3
    m.set_params(p)
    m.set_kind(nn.KIND_CORRECTION, cs)
    _, out, _, _ = m.forward_backward(ravg.astype(np.float32), np.zeros((nc,) * 3, np.float32),
                                      np.zeros(4))
    if kind == "constant":
        assert np.all(out == np.float32(arg))
    sc = cs * np.sqrt(np.mean(ravg ** 2))
    phi_c = phi_w.copy()
    pe = interpolate(cubic_weights(n, nc), out.astype(np.float64))
    phi_c[1:-1, 1:-1, 1:-1] += sc * pe[1:-1, 1:-1, 1:-1]
    want = np.linalg.norm(interior_residual(phi_c, f, h)) / np.linalg.norm(rw)

    g = solver(1e-8)
    g.set_mode(H.POISSON_NN_FP32_DC)
    g.set_nn_correction(m)
    g.set_phi(phi_w)
    assert g.solve_potential(rho) == 0
    s = g.s
    assert s.last_warm_rel_residual > 0 and s.last_corr_rel_residual > 0
    assert s.last_corr_rel_residual / s.last_warm_rel_residual == pytest.approx(want, rel=tol)
    assert s.last_guess == (2 if s.last_corr_rel_residual < s.last_warm_rel_residual else 0)
    assert s.last_rel_residual <= 1e-8
    assert rel_l2(g.phi(), phi_ref) < 1e-7


def test_no_correction_once_the_start_is_converged():
    """A start already within the tolerance is kept as it is."""
    n = 64
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = 1e-3 + np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / 0.02)
    g = H.Gravity(n, n, n, G=1.0)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_params(1.5, 400, 1e-8)
    g.set_mode(H.POISSON_FP32_DC)
    g.set_dc_inner(H.DC_INNER_MULTIGRID)
    assert g.solve_potential(rho) == 0
    m = nn.Model(seed=0)
    m.set_kind(nn.KIND_CORRECTION, 1e-3)
    g.set_mode(H.POISSON_NN_FP32_DC)
    g.set_nn_correction(m)
    assert g.solve_potential(rho) == 0
    assert g.s.last_corr_rel_residual == -1.0 and g.s.last_guess == -1
    assert g.diagnostics()[0] == 0


def explosion_sweeps(n, seed, steps, weights, inner):
    """Per-step inner sweeps (or V-cycles) and start used by the FP32 defect
    correction of an explosion run, with the start correction attached when
    `weights` is given."""
    U = ics.explosion_n3(n, seed=seed)
    s = H.Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
    g = s.gravity
    g.set_lmax(6)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_params(1.9 if inner == H.DC_INNER_SOR else 1.5, 400000, 1e-8)
    g.set_mode(H.POISSON_NN_FP32_DC if weights else H.POISSON_FP32_DC)
    g.set_dc_inner(inner)
    if weights:
        m = nn.Model(path=weights)
        g.set_nn_correction(m)
    s.set_state(U)
    its, used = [], []
    for _ in range(steps):
        assert s.step() == 0
        assert g.s.last_status == 0
        its.append(g.s.last_iterations)
        used.append(g.s.last_guess)
    s.close()
    return np.array(its), np.array(used)


@pytest.mark.skipif(not os.path.exists(nn.CORRECTION_WEIGHTS),
                    reason="no trained start-correction weights")
def test_trained_correction_on_a_held_out_run():
    """The committed start correction on the 128^3 explosion with the held-out
    seed 12345: the corrected start wins on most warm steps and the FP32
    defect correction with SOR sweeps needs fewer of them."""
    m = nn.Model(path=nn.CORRECTION_WEIGHTS)
    assert m.kind()[0] == nn.KIND_CORRECTION and m.kind()[1] > 0
    m.close()
    base, _ = explosion_sweeps(128, 12345, 24, None, H.DC_INNER_SOR)
    corr, used = explosion_sweeps(128, 12345, 24, nn.CORRECTION_WEIGHTS, H.DC_INNER_SOR)
    warm = slice(4, None)
    assert np.mean(used[warm] == 2) >= 0.75, used
    assert corr[warm].sum() < 0.95 * base[warm].sum(), (corr, base)


@pytest.mark.skipif(not os.path.exists(nn.CORRECTION_WEIGHTS),
                    reason="no trained start-correction weights")
def test_trained_correction_at_256_with_multigrid():
    """The committed start correction on the 256^3 stellar explosion (seed
    12345, held out of training) with multigrid inner solves: the corrected
    start wins the warm steps and the run needs fewer V-cycles."""
    base, _ = explosion_sweeps(256, 12345, 12, None, H.DC_INNER_MULTIGRID)
    corr, used = explosion_sweeps(256, 12345, 12, nn.CORRECTION_WEIGHTS, H.DC_INNER_MULTIGRID)
    warm = slice(1, None)
    assert np.mean(used[warm] == 2) >= 0.9, used
    assert corr[warm].sum() < base[warm].sum(), (corr, base)


def test_potential_guess_skipped_while_it_cannot_win():
    """The potential guess is tried on a cold start and again after a
    network is attached, and skipped while the best start has under a
    quarter of the residual the guess had when last tried."""
    n = 64
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = 1e-3 + np.exp(-((x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.5) ** 2) / 0.02)
    g = H.Gravity(n, n, n, G=1.0)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_params(1.5, 400, 1e-8)
    g.set_mode(H.POISSON_NN_FP32_DC)
    g.set_dc_inner(H.DC_INNER_MULTIGRID)
    m = nn.Model(seed=4)
    g.set_nn_model(m)
    assert g.solve_potential(rho) == 0
    tried = g.s.last_guess_rel_residual
    assert tried > 0 and g.s.last_guess in (0, 1)
    # the converged potential as the start: far under a quarter of the guess's
    assert g.solve_potential(rho) == 0
    assert g.s.last_guess_rel_residual == -1.0 and g.s.last_guess == 0
    assert 4.0 * g.s.last_warm_rel_residual < tried
    assert g.diagnostics()[0] == 0
    g.set_nn_model(m)  # attaching a network makes the next solve try it
    assert g.solve_potential(rho) == 0
    # the guess's input statistics are summed with atomics, so the same
    # density gives the same guess to about 1e-5 (measured 4.5e-5 here)
    assert g.s.last_guess_rel_residual == pytest.approx(tried, rel=1e-3)
    assert g.s.last_guess == 0
