"""The learned Poisson guess on the device: forward and
backward passes against the torch CPU reference (tests/nn_reference.py),
training, and the guess-or-warm-start rule of POISSON_NN_FP32_DC."""
import numpy as np
import pytest

import hydrocode_astro1_amd as H
from hydrocode_astro1_amd import ics, nn
import nn_reference as R

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def sample(n, seed):
    rng = np.random.default_rng(seed)
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = (1e-3 + np.exp(-((x - 0.45) ** 2 + (y - 0.55) ** 2 + (z - 0.5) ** 2) / 0.02)
           * (1.0 + 0.2 * rng.standard_normal((n, n, n)) ** 2)).astype(np.float32)
    mom4 = np.array([1.0, 0.47, 0.53, 0.5])
    target = (R.mono_field(n, mom4) + 0.05 * np.sin(6 * x) * np.cos(4 * y)).astype(np.float32)
    return rho, target, mom4


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_backward_match_torch(seed):
    n = 32
    m = nn.Model(seed=seed)
    # non-trivial last layer so that every layer carries signal
    p = m.get_params()
    rng = np.random.default_rng(100 + seed)
    p += (0.02 * rng.standard_normal(p.size)).astype(np.float32) * (p != 0)
    m.set_params(p)
    rho, target, mom4 = sample(n, seed)
    loss, out, feat, grads = m.forward_backward(rho, target, mom4)

    # features: normalised log-density and 0.1 x monopole, BF16
    l10 = np.log10(np.maximum(rho.astype(np.float64), 1e-12))
    f0 = (l10 - l10.mean()) / l10.std()
    assert np.max(np.abs(feat[..., 0] - f0)) < 2e-2 * np.max(np.abs(f0))
    mono = R.mono_field(n, mom4)
    assert np.max(np.abs(feat[..., 1] - 0.1 * mono)) <= 2 ** -8 * np.max(np.abs(0.1 * mono))

    lp = m.layer_params(p)
    # the reference rounds gradients to BF16 where the device stores them;
    # what is left is FP32 sums taken in another order that land on the
    # other side of a BF16 rounding boundary (measured: output 4e-4, weight
    # and bias gradients at most 1.1e-2 relative L2; without the gradient
    # rounding the coarsest layers differ by 6%)
    rloss, rout, rgrads = R.loss_and_grads(lp, feat, mono, target, grad_rounding=True)
    assert rel_l2(out, rout) < 1e-2
    assert loss == pytest.approx(rloss, rel=3e-3)
    glp = m.layer_params(grads)
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(glp, rgrads)):
        assert rel_l2(gw, rw) < 3e-2, f"layer {l} weight gradient"
        assert rel_l2(gb, rb) < 3e-2, f"layer {l} bias gradient"
        cos = float(np.dot(gw.ravel(), rw.ravel()) /
                    (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        assert cos > 0.9995, f"layer {l} weight gradient direction"


def test_device_gradients_match_finite_differences():
    """Central differences of the device's own loss against its gradient,
    layer by layer: each layer's parameters move along its gradient by 0.3%
    of their RMS, so the BF16 weights do change while the ReLU pattern
    mostly does not (along random +-1 directions the projection g.d cancels
    and BF16 rounding of the activations swamps it; beyond a few percent the
    kinks do).  Measured ratios 0.992-1.024."""
    n = 32
    m = nn.Model(seed=4)
    p = m.get_params()
    rng = np.random.default_rng(7)
    p += (0.02 * rng.standard_normal(p.size)).astype(np.float32) * (p != 0)
    m.set_params(p)
    rho, target, mom4 = sample(n, 2)
    _, _, _, grads = m.forward_backward(rho, target, mom4)
    for l, (cin, cout, _) in enumerate(nn.LAYERS):
        wo = m.L.nn_model_param_offset(m.h, l, 0)
        bo = m.L.nn_model_param_offset(m.h, l, 1)
        sl = np.arange(wo, bo + cout)
        live = sl[p[sl] != 0.0]  # padded input channels stay zero
        g = grads[live].astype(np.float64)
        d = np.zeros_like(p)
        d[live] = (0.003 * np.sqrt(np.mean(p[live].astype(np.float64) ** 2)) * g /
                   np.sqrt(np.mean(g ** 2))).astype(np.float32)
        m.set_params(p + d)
        lp, _, _, _ = m.forward_backward(rho, target, mom4)
        m.set_params(p - d)
        lm, _, _, _ = m.forward_backward(rho, target, mom4)
        fd = (lp - lm) / 2.0
        an = float(np.dot(grads.astype(np.float64), d.astype(np.float64)))
        print(f"layer {l}: finite difference {fd:.5e}, gradient {an:.5e}, ratio {fd / an:.4f}")
        assert fd == pytest.approx(an, rel=0.05), f"layer {l}"
    m.set_params(p)


@pytest.fixture(scope="module")
def small_set():
    ds = nn.build_dataset(32, 12, first_seed=500, log=None)
    return ds


def test_dataset_targets_solve_poisson(small_set):
    """Stored potential = the solver's potential in units of G M / L: check
    the discrete Poisson equation on the interior against the stored
    density (FP32 storage limits the match)."""
    rho, phi, mom = small_set.get(3)
    n = 32
    h = 1.0 / n
    assert mom[0] == pytest.approx(float(np.sum(rho.astype(np.float64))) * h ** 3, rel=1e-5)
    p = phi.astype(np.float64) * mom[0]  # G = L = 1
    lap = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1]
           + p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6 * p[1:-1, 1:-1, 1:-1]) / h ** 2
    f = 4 * np.pi * rho[1:-1, 1:-1, 1:-1].astype(np.float64)
    assert rel_l2(lap, f) < 1e-3


def test_training_reduces_loss(small_set):
    m = nn.Model(seed=3)
    m0 = np.zeros(1, dtype=np.float32)
    L = H.lib()
    assert L.nn_evaluate(m.h, small_set.h, 0, 12, m0.ctypes.data_as(nn._fp)) == 0
    losses, vmse = nn.train(m, small_set, epochs=6, batch=2, lr=3e-3, val=0, log=None)
    m1 = np.zeros(1, dtype=np.float32)
    assert L.nn_evaluate(m.h, small_set.h, 0, 12, m1.ctypes.data_as(nn._fp)) == 0
    assert np.all(np.isfinite(losses))
    assert m1[0] < 0.8 * m0[0]


def test_save_load_roundtrip(tmp_path):
    a = nn.Model(seed=5)
    f = tmp_path / "w.bin"
    a.save(str(f))
    b = nn.Model(seed=6, path=str(f))
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    with pytest.raises(RuntimeError):
        b.load(str(tmp_path / "missing.bin"))


@pytest.mark.parametrize("warm", [False, True])
def test_guess_or_warm_start(warm):
    """POISSON_NN_FP32_DC keeps whichever start has the smaller residual and
    then converges like POISSON_FP32_DC to the FP64 solution."""
    n = 64
    X = (np.arange(n) + 0.5) / n
    x, y, z = np.meshgrid(X, X, X, indexing="ij")
    rho = 1e-3 + np.exp(-((x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.55) ** 2) / 0.01)
    ref = H.Gravity(n, n, n, G=1.0)
    ref.set_lmax(4)
    ref.set_criterion(H.SOR_CRIT_REL_RESID)
    ref.set_params(1.9, 200000, 1e-10)
    assert ref.solve_potential(rho) == 0
    phi_ref = ref.phi()

    g = H.Gravity(n, n, n, G=1.0)
    g.set_lmax(4)
    g.set_criterion(H.SOR_CRIT_REL_RESID)
    g.set_params(1.9, 200000, 1e-8)
    m = nn.Model(seed=1)
    g.set_nn_model(m)
    g.set_mode(H.POISSON_NN_FP32_DC)
    if warm:
        g.set_phi(phi_ref)
    assert g.solve_potential(rho) == 0
    s = g.s
    assert s.last_guess in (0, 1)
    assert s.last_guess_rel_residual > 0 and s.last_warm_rel_residual > 0
    chose_nn = s.last_guess_rel_residual < s.last_warm_rel_residual
    assert s.last_guess == int(chose_nn)
    if warm:
        assert s.last_guess == 0  # the converged potential beats any guess
    assert s.last_rel_residual <= 1e-8
    assert rel_l2(g.phi(), phi_ref) < 1e-7


def test_deck_attaches_the_network(tmp_path):
    """poisson_mode = nn_fp32dc with nn_guess = <weights>: the solver made
    from the deck owns a loaded network and its first step keeps the better
    start and converges."""
    n = 32
    deck = tmp_path / "deck.txt"
    deck.write_text(f"nx = {n}\nny = {n}\nnz = {n}\nG_const = 1.0\ngamma = 2.0\n"
                    f"sor_tol = 1e-8\nsor_criterion = rel_resid\npoisson_mode = nn_fp32dc\n"
                    f"nn_guess = {nn.DEFAULT_WEIGHTS}\noutput_enable = false\n")
    L = H.lib()
    d = L.input_deck_create(str(deck).encode())
    h = L.solver3d_create_from_deck(d)
    L.input_deck_destroy(d)
    assert h
    s = H.Solver(0, 0, 0, handle=h)
    g = s.s.gravity_solver.contents
    assert g.poisson_mode == H.POISSON_NN_FP32_DC and g.nn_model
    X, Y, Z = ics.centres((n, n, n))
    rho = 1e-3 + np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) / 0.01)
    U = ics.prim_to_cons(rho, 0 * X, 0 * X, 0 * X, 0.1 * rho ** 2, 2.0)
    s.set_state(U)
    assert s.step() == 0
    assert g.last_status == 0 and g.last_guess in (0, 1)
    assert g.last_rel_residual <= 1e-8
    s.close()
