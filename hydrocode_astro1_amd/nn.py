"""Learned initial guess of the Poisson solve: the network, its training set
and its training loop, all on the device through libhydro_mi355x.so.

The reference has no counterpart (its SOR starts from the previous potential
or from zero, gravity_solver_3d.c:87 and :566).  The network is a 4-level 3D
U-Net with 3x3x3 convolutions (65k parameters) that maps the normalised
log-density and the softened monopole potential about the centre of mass to
the potential in units of G M / L, as the monopole plus a learned
correction.  Convolutions run on MFMA in BF16 with FP32 accumulation; the
master weights and Adam state are FP32.

Training data (the 128^3 learned-guess workload): fields of 1-8 blobs
(Gaussian or n=1 polytrope; masses log-uniform in [0.1, 10], widths uniform
in [0.03, 0.15], centres uniform in [0.2, 0.8]^3) on a floor density, times a
lognormal k^-2 turbulence factor; each potential is this library's own
isolated solve to a relative residual of 1e-8.

    python -m hydrocode_astro1_amd.nn --fields 2000 --epochs 4 --out weights.bin

The same network has a second kind, the start correction (kind 1).  Its
input is the residual r = f - L phi of a warm start (the previous step's
potential with the new density's boundary values) averaged onto 64^3, its
output the correction e of the start, in units of C L^2 s (s the root mean
square of the averaged r, L the box edge, C a constant stored with the
weights).  Training pairs come from explosion runs (the n=3 polytrope of
ics.explosion_n3, seeds 1-8 at 128^3 and 21-22 at 256^3) whose every step is
solved to a relative residual of 1e-11; the held-out seed 12345 is the
stellar-explosion workload.

    python -m hydrocode_astro1_amd.nn --kind correction --out corr.bin
"""
import argparse
import ctypes as C
import math
import os
import sys
import time

import numpy as np

from . import (DC_INNER_MULTIGRID, DC_INNER_SOR, GRAV_BC_ISOLATED, POISSON_FP32_DC,
               SOR_CRIT_REL_RESID, Gravity, _check, lib)

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "nn_guess_128.bin")
CORRECTION_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                  "nn_correction_64.bin")
KIND_GUESS, KIND_CORRECTION = 0, 1
# (cin, cout, nearest-upsampled input) of the 12 convolutions, in order
LAYERS = [(2, 8, 0), (8, 8, 0), (8, 12, 0), (12, 12, 0), (12, 16, 0), (16, 16, 0),
          (16, 24, 0), (24, 24, 0), (24, 16, 1), (16, 12, 1), (12, 8, 1), (8, 1, 0)]


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def pad4(c):
    return (c + 3) & ~3


class Model:
    """The guess network (nn_model_create); weights loaded from `path` when
    given."""

    def __init__(self, seed=0, path=None):
        self.L = lib()
        self.h = self.L.nn_model_create(seed)
        if not self.h:
            raise RuntimeError("nn_model_create failed")
        self.nparam = self.L.nn_model_param_len(self.h)
        if path is not None:
            self.load(path)

    def close(self):
        if getattr(self, "h", None):
            self.L.nn_model_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def num_params(self):
        return self.L.nn_model_num_params(self.h)

    def load(self, path):
        if self.L.nn_model_load(self.h, os.fsencode(path)) != 0:
            raise RuntimeError(f"cannot load network weights from {path}")

    def save(self, path):
        if self.L.nn_model_save(self.h, os.fsencode(path)) != 0:
            raise RuntimeError(f"cannot save network weights to {path}")

    def set_kind(self, kind, cscale=1.0):
        """KIND_GUESS, or KIND_CORRECTION with its output scale C."""
        if self.L.nn_model_set_kind(self.h, int(kind), float(cscale)) != 0:
            raise ValueError(f"bad network kind {kind} or scale {cscale}")

    def kind(self):
        """(kind, C)."""
        c = C.c_double(0.0)
        k = self.L.nn_model_kind(self.h, C.byref(c))
        return k, c.value

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        _check(self.L.nn_model_get_params(self.h, _f(p)), "nn_model_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _check(self.L.nn_model_set_params(self.h, _f(p)), "nn_model_set_params")

    def layer_params(self, p):
        """Split a packed parameter vector into [(W[27][cin][cout], b[cout])]
        with the padded input channels dropped."""
        out = []
        for l, (cin, cout, _) in enumerate(LAYERS):
            wo = self.L.nn_model_param_offset(self.h, l, 0)
            bo = self.L.nn_model_param_offset(self.h, l, 1)
            w = p[wo:wo + 27 * pad4(cin) * cout].reshape(27, pad4(cin), cout)[:, :cin, :]
            out.append((w, p[bo:bo + cout]))
        return out

    def forward_backward(self, rho, target, mom4, box_len=1.0):
        """One pass on host data: (loss, output, features[n^3][2], gradients)."""
        rho = np.ascontiguousarray(rho, dtype=np.float32)
        n = rho.shape[0]
        target = np.ascontiguousarray(target, dtype=np.float32)
        mom = np.ascontiguousarray(mom4, dtype=np.float64)
        loss = np.zeros(1, dtype=np.float32)
        out = np.zeros(rho.shape, dtype=np.float32)
        feat = np.zeros(rho.shape + (2,), dtype=np.float32)
        grads = np.zeros(self.nparam, dtype=np.float32)
        rc = self.L.nn_model_forward_backward(self.h, _f(rho), _f(target), n, _d(mom),
                                              float(box_len), _f(loss), _f(out), _f(feat),
                                              _f(grads))
        if rc != 0:
            raise RuntimeError(f"nn_model_forward_backward failed with {rc}")
        return float(loss[0]), out, feat, grads


class Dataset:
    """Device-resident training set (nn_dataset_create)."""

    def __init__(self, n, count):
        self.L = lib()
        self.n, self.count = n, count
        self.h = self.L.nn_dataset_create(n, count)
        if not self.h:
            raise RuntimeError("nn_dataset_create failed")

    def close(self):
        if getattr(self, "h", None):
            self.L.nn_dataset_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fill(self, index, params, grav):
        blob, mode, floor_rho, sigma = params
        blob = np.ascontiguousarray(blob, dtype=np.float64).reshape(-1)
        mode = np.ascontiguousarray(mode, dtype=np.float64).reshape(-1)
        return _check(self.L.nn_dataset_fill(self.h, index, _d(blob), blob.size // 6, _d(mode),
                                             mode.size // 5, float(floor_rho), float(sigma),
                                             grav.h), "nn_dataset_fill")

    def set(self, index, inp, target):
        """Field `index` from host arrays: the network input and its target."""
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        assert inp.shape == target.shape == (self.n,) * 3
        _check(self.L.nn_dataset_set(self.h, index, _f(inp), _f(target)), "nn_dataset_set")

    def get(self, index):
        n = self.n
        rho = np.zeros((n, n, n), dtype=np.float32)
        phi = np.zeros((n, n, n), dtype=np.float32)
        mom = np.zeros(4)
        _check(self.L.nn_dataset_get(self.h, index, _f(rho), _f(phi), _d(mom)), "nn_dataset_get")
        return rho, phi, mom


def field_params(seed):
    """Parameters of training field `seed` for blob_field_kernel:
    (blobs [nb][cx cy cz mass width type], modes [nm][kx ky kz phase amp],
    floor density, sigma of the lognormal factor)."""
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(1, 9))
    blobs = np.zeros((nb, 6))
    blobs[:, 0:3] = rng.uniform(0.2, 0.8, size=(nb, 3))
    blobs[:, 3] = 10.0 ** rng.uniform(-1.0, 1.0, size=nb)
    blobs[:, 4] = rng.uniform(0.03, 0.15, size=nb)
    blobs[:, 5] = rng.integers(0, 2, size=nb)
    # k^-2 power: amplitude ~ 1/|k| per mode, scaled to the requested sigma
    nm = 48
    k = rng.integers(-8, 9, size=(nm, 3)).astype(np.float64)
    k[np.all(k == 0, axis=1)] = (1.0, 0.0, 0.0)
    kk = np.linalg.norm(k, axis=1)
    amp = 1.0 / kk
    sigma = rng.uniform(0.05, 0.5)
    amp *= sigma / math.sqrt(0.5 * np.sum(amp ** 2))
    modes = np.column_stack([k, rng.uniform(0, 2 * math.pi, size=nm), amp])
    floor_rho = 10.0 ** rng.uniform(-4.0, -2.0)
    return blobs, modes, floor_rho, sigma


def make_solver(n, omega=1.9, tol=1e-8, lmax=6, inner=None):
    """Isolated n^3 solver on the unit box used to produce training targets:
    FP32 defect correction to a relative residual `tol`, its corrections
    solved by multigrid V-cycles where the grid allows it (even n >= 8),
    else by red-black SOR sweeps with factor `omega`."""
    if inner is None:
        inner = DC_INNER_MULTIGRID if n % 2 == 0 and n >= 8 else DC_INNER_SOR
    g = Gravity(n, n, n, G=1.0)
    g.set_boundary(GRAV_BC_ISOLATED)
    g.set_lmax(lmax)
    g.set_params(omega, 400000, tol)
    g.set_mode(POISSON_FP32_DC)
    g.set_dc_inner(inner)
    g.set_criterion(SOR_CRIT_REL_RESID)
    return g


def build_dataset(n, count, first_seed=0, omega=1.9, tol=1e-8, log=print, inner=None):
    """Generate and solve fields first_seed .. first_seed + count - 1."""
    ds = Dataset(n, count)
    grav = make_solver(n, omega, tol, inner=inner)
    t0 = time.perf_counter()
    bad = 0
    for i in range(count):
        st = ds.fill(i, field_params(first_seed + i), grav)
        bad += st != 0
        if log and (i + 1) % max(1, count // 10) == 0:
            log(f"[nn] fields {i + 1}/{count} solved in {time.perf_counter() - t0:.1f}s "
                f"(last: {grav.s.last_iterations} sweeps or V-cycles, rel. residual "
                f"{grav.s.last_rel_residual:.2e}, unconverged so far {bad})")
    grav.close()
    return ds


def train(model, ds, epochs=4, batch=4, lr=2e-3, val=0, seed=0, log=print, decay=0.5):
    """Adam over shuffled mini-batches of fields [0, count - val), learning
    rate lr * decay^epoch; returns (per-step losses, validation MSE on the
    last `val` fields or None)."""
    L = lib()
    rng = np.random.default_rng(seed)
    ntrain = ds.count - val
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(ntrain).astype(np.int32)
        steps = ntrain // batch
        order = np.ascontiguousarray(order[:steps * batch])
        lr_ep = lr * (decay ** ep)
        ll = np.zeros(max(steps, 1), dtype=np.float32)
        rc = L.nn_train(model.h, ds.h, order.ctypes.data_as(_ip), steps, batch, lr_ep, _f(ll))
        if rc != 0:
            raise RuntimeError(f"nn_train failed with {rc}")
        losses.extend(ll[:steps].tolist())
        if log:
            log(f"[nn] epoch {ep + 1}/{epochs}: mean loss {float(np.mean(ll[:steps])):.3e} "
                f"({time.perf_counter() - t0:.1f}s)")
    vmse = None
    if val > 0:
        m = np.zeros(1, dtype=np.float32)
        rc = L.nn_evaluate(model.h, ds.h, ntrain, val, _f(m))
        if rc != 0:
            raise RuntimeError(f"nn_evaluate failed with {rc}")
        vmse = float(m[0])
    return losses, vmse


# ---- start correction ------------------------------------------------------
def average(a, f):
    """Mean over f^3 blocks (n^3 -> (n/f)^3)."""
    m = a.shape[0] // f
    return a.reshape(m, f, m, f, m, f).mean(axis=(1, 3, 5))


def zero_boundary(e):
    e[0] = e[-1] = 0.0
    e[:, 0] = e[:, -1] = 0.0
    e[:, :, 0] = e[:, :, -1] = 0.0
    return e


def laplacian(e, h):
    """7-point Laplacian on the interior cells (boundary cells 0)."""
    r = np.zeros_like(e)
    c = e[1:-1, 1:-1, 1:-1]
    r[1:-1, 1:-1, 1:-1] = ((e[2:, 1:-1, 1:-1] + e[:-2, 1:-1, 1:-1] - 2.0 * c) +
                           (e[1:-1, 2:, 1:-1] + e[1:-1, :-2, 1:-1] - 2.0 * c) +
                           (e[1:-1, 1:-1, 2:] + e[1:-1, 1:-1, :-2] - 2.0 * c)) / (h * h)
    return r


def correction_pairs(n, seed, steps, nc=64, tol=1e-11, log=print):
    """(averaged residual, averaged correction) pairs of `steps` consecutive
    self-gravity steps of the explosion at n^3 (unit box, G = 1, outflow
    boundaries): each step's potential is solved to the relative residual
    `tol`, e = phi_new - phi_old with the boundary cells zeroed (the start
    takes the new boundary values), r = L e = f - L phi_old on the interior."""
    from . import BC_OUTFLOW, Solver, ics
    U = ics.explosion_n3(n, seed=seed)
    s = Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(BC_OUTFLOW,) * 6)
    g = s.gravity
    g.set_lmax(6)
    g.set_criterion(SOR_CRIT_REL_RESID)
    g.set_mode(POISSON_FP32_DC)
    g.set_dc_inner(DC_INNER_MULTIGRID)
    g.set_params(1.5, 400, tol)
    s.set_state(U)
    del U
    f = n // nc
    h = 1.0 / n
    prev = None
    R, E = [], []
    t0 = time.perf_counter()
    for k in range(steps + 1):
        s.step()
        lib().gravity_sync_to_host(g.h)
        phi = g.phi().copy()
        if prev is not None:
            e = zero_boundary(phi - prev)
            R.append(average(laplacian(e, h), f).astype(np.float32))
            E.append(average(e, f).astype(np.float32))
        prev = phi
        if log and (k + 1) % 40 == 0:
            log(f"[nn] {n}^3 seed {seed}: step {k + 1}/{steps + 1} "
                f"({time.perf_counter() - t0:.1f}s)")
    s.close()
    return R, E


def correction_scale(R, E, box_len=1.0):
    """C: the median over pairs of rms(e) / (rms(r) L^2)."""
    sr = np.array([np.sqrt(np.mean(np.square(r, dtype=np.float64))) for r in R])
    se = np.array([np.sqrt(np.mean(np.square(e, dtype=np.float64))) for e in E])
    ok = sr > 0
    return float(np.median(se[ok] / (sr[ok] * box_len * box_len)))


def correction_dataset(R, E, cscale, box_len=1.0):
    """Device dataset of the pairs with pairs of a zero residual dropped:
    input r, target e / (C L^2 s)."""
    keep = [i for i in range(len(R)) if np.any(R[i])]
    ds = Dataset(R[0].shape[0], len(keep))
    for j, i in enumerate(keep):
        s = math.sqrt(float(np.mean(np.square(R[i], dtype=np.float64))))
        ds.set(j, R[i], E[i] / (cscale * box_len * box_len * s))
    return ds


def train_correction(runs, epochs=40, batch=8, lr=2e-3, decay=0.92, val_frac=0.1, seed=0,
                     log=print):
    """Pairs of the runs [(n, seed, steps)], then the start-correction
    network trained on them; returns (model, validation MSE)."""
    R, E = [], []
    for n, sd, steps in runs:
        r, e = correction_pairs(n, sd, steps, log=log)
        R += r
        E += e
        if log:
            log(f"[nn] pairs so far {len(R)}")
    C_ = correction_scale(R, E)
    perm = np.random.default_rng(seed).permutation(len(R))
    R = [R[i] for i in perm]
    E = [E[i] for i in perm]
    ds = correction_dataset(R, E, C_)
    val = max(1, int(round(val_frac * ds.count)))
    model = Model(seed=seed)
    model.set_kind(KIND_CORRECTION, C_)
    if log:
        log(f"[nn] {ds.count} pairs, C = {C_:.4e}")
    _, vmse = train(model, ds, epochs, batch, lr, val=val, seed=seed, log=log, decay=decay)
    ds.close()
    return model, vmse


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--fields", type=int, default=2000)
    ap.add_argument("--val", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=None, help="4 (guess), 40 (correction)")
    ap.add_argument("--batch", type=int, default=None, help="4 (guess), 8 (correction)")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--decay", type=float, default=None,
                    help="learning-rate factor per epoch: 0.5 (guess), 0.92 (correction)")
    ap.add_argument("--omega", type=float, default=1.9, help="SOR factor of the sor inner solve")
    ap.add_argument("--inner", choices=("auto", "sor", "multigrid"), default="auto",
                    help="inner solve of the training targets' defect correction")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--kind", choices=("guess", "correction"), default="guess")
    ap.add_argument("--runs", default="128:1:160,128:2:160,128:3:160,128:4:160,128:5:160,"
                    "128:6:160,128:7:160,128:8:160,256:21:130,256:22:130",
                    help="start correction: explosion runs n:seed:steps")
    a = ap.parse_args(argv)
    corr = a.kind == "correction"
    a.epochs = a.epochs or (40 if corr else 4)
    a.batch = a.batch or (8 if corr else 4)
    a.decay = a.decay or (0.92 if corr else 0.5)
    if corr:
        runs = [tuple(int(v) for v in r.split(":")) for r in a.runs.split(",") if r]
        model, vmse = train_correction(runs, epochs=a.epochs, batch=a.batch, lr=a.lr,
                                       decay=a.decay, seed=a.seed)
        out = a.out or CORRECTION_WEIGHTS
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        model.save(out)
        print(f"[nn] start correction: validation MSE {vmse}, C = {model.kind()[1]:.4e}, "
              f"saved {out}", file=sys.stderr)
        return
    a.out = a.out or DEFAULT_WEIGHTS
    inner = {"auto": None, "sor": DC_INNER_SOR, "multigrid": DC_INNER_MULTIGRID}[a.inner]
    ds = build_dataset(a.n, a.fields, omega=a.omega, inner=inner)
    model = Model(seed=a.seed)
    losses, vmse = train(model, ds, a.epochs, a.batch, a.lr, val=a.val, seed=a.seed,
                         decay=a.decay)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    print(f"[nn] {model.num_params()} parameters, final loss {losses[-1]:.3e}, "
          f"validation MSE {vmse}, saved {a.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
