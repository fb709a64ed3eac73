"""Initial conditions for the benchmark workloads, generated on the host.

None of these problems has a generator in the reference
(hydro_solver_3d.c:1383-1515 and :1864-2133 cover other set-ups), so they
are built here from their definitions and handed to the solver (and to the
CPU oracle) as the same conserved-state array:

  sod            3D Sod tube along x, 64x16x16, gamma 1.4
  polytrope_n1   n=1 polytrope rho_c sin(pi r/R)/(pi r/R), R 0.35, gamma 2
  explosion_n3   Lane-Emden n=3 star, R 0.3, lognormal k^-2 perturbations,
                 thermal energy 10|W| in the central 4 cells
  blobs          1-8 Gaussian / n=1 polytrope blobs plus lognormal
                 turbulence (densities for the learned initial guess)

Arrays are (nx, ny, nz, 5) float64 in the reference's index order
(IDX3D, hydro_solver_3d.h:12) with fields rho, rho u, rho v, rho w, E.
"""
import numpy as np

IC_MAGIC = b"HYDIC001"


def centres(n3, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    nx, ny, nz = n3
    x0, x1, y0, y1, z0, z1 = box
    x = x0 + (np.arange(nx) + 0.5) * ((x1 - x0) / nx)
    y = y0 + (np.arange(ny) + 0.5) * ((y1 - y0) / ny)
    z = z0 + (np.arange(nz) + 0.5) * ((z1 - z0) / nz)
    return np.meshgrid(x, y, z, indexing="ij", sparse=True)


def prim_to_cons(rho, u, v, w, p, gamma):
    """prim_to_cons_3d (hydro_solver_3d.c:177-189) on arrays."""
    shape = np.broadcast(rho, u, v, w, p).shape
    U = np.empty(shape + (5,))
    U[..., 0] = rho
    U[..., 1] = rho * u
    U[..., 2] = rho * v
    U[..., 3] = rho * w
    ke = 0.5 * rho * (u * u + v * v + w * w)
    U[..., 4] = p / (gamma - 1.0) + ke
    return U


def sod(nx=64, ny=16, nz=16, gamma=1.4):
    """Left half rho=1 p=1, right half rho=0.125 p=0.1, at rest."""
    X, _, _ = centres((nx, ny, nz))
    left = np.broadcast_to(X < 0.5, (nx, ny, nz))
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros((nx, ny, nz))
    return prim_to_cons(rho, z, z, z, p, gamma)


def polytrope_n1(n=128, R=0.35, rho_c=1.0, floor=1e-6, G=1.0, gamma=2.0):
    """n=1 polytrope in hydrostatic balance: p = K rho^2 with K = 2 G R^2 / pi
    (Lane-Emden n=1: R = pi sqrt(K / (2 pi G)))."""
    X, Y, Z = centres((n, n, n))
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    x = np.pi * r / R
    with np.errstate(invalid="ignore", divide="ignore"):
        prof = np.where(x > 0, np.sin(x) / np.where(x > 0, x, 1.0), 1.0)
    rho = np.where(r < R, rho_c * prof, floor)
    rho = np.maximum(rho, floor)
    K = 2.0 * G * R * R / np.pi
    p = K * rho * rho
    z = np.zeros_like(rho)
    return prim_to_cons(rho, z, z, z, p, gamma)


def lane_emden(nidx, dxi=1e-4):
    """theta(xi) of the Lane-Emden equation by RK4; returns (xi, theta,
    xi1, xi1^2 |theta'(xi1)|)."""
    def f(xi, y):
        th, dth = y
        return np.array([dth, -max(th, 0.0) ** nidx - 2.0 * dth / xi])

    xi = 1e-6
    y = np.array([1.0 - xi * xi / 6.0, -xi / 3.0])
    xs, ts = [0.0], [1.0]
    while y[0] > 0.0:
        k1 = f(xi, y)
        k2 = f(xi + 0.5 * dxi, y + 0.5 * dxi * k1)
        k3 = f(xi + 0.5 * dxi, y + 0.5 * dxi * k2)
        k4 = f(xi + dxi, y + dxi * k3)
        yn = y + dxi / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        if yn[0] <= 0.0:  # linear interpolation to the surface
            frac = y[0] / (y[0] - yn[0])
            xi1 = xi + frac * dxi
            dth1 = y[1] + frac * (yn[1] - y[1])
            xs.append(xi1)
            ts.append(0.0)
            return np.array(xs), np.array(ts), xi1, xi1 * xi1 * abs(dth1)
        xi += dxi
        y = yn
        xs.append(xi)
        ts.append(y[0])
    raise RuntimeError("Lane-Emden integration did not reach the surface")


def gaussian_field(n, slope, sigma, seed):
    """Isotropic Gaussian random field with P(k) ~ k^slope, zero mean, std
    sigma, periodic on an n^3 grid."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n, n, n))
    fk = np.fft.rfftn(noise)
    k = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    amp[kk > 0] = kk[kk > 0] ** (slope / 2.0)
    field = np.fft.irfftn(fk * amp, s=(n, n, n), axes=(0, 1, 2))
    field -= field.mean()
    field *= sigma / field.std()
    return field


def explosion_n3(n=256, R=0.3, rho_c=1.0, env=1e-8, G=1.0, gamma=5.0 / 3.0, sigma=0.01,
                 seed=12345, e_factor=10.0, r_cells=4.0):
    """n=3 polytrope with seeded lognormal density perturbations and a
    central thermal energy deposit of e_factor * |W|."""
    xi, th, xi1, mu = lane_emden(3)
    alpha = R / xi1
    K = np.pi * G * alpha * alpha * rho_c ** (2.0 / 3.0)
    M = 4.0 * np.pi * rho_c * alpha ** 3 * mu
    W = -1.5 * G * M * M / R
    X, Y, Z = centres((n, n, n))
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    theta = np.interp(r / alpha, xi, th, right=0.0)
    rho = np.where(r < R, rho_c * theta ** 3, env)
    rho = np.maximum(rho, env)
    p = K * rho ** (4.0 / 3.0)
    if sigma > 0:
        d = gaussian_field(n, -2.0, sigma, seed)
        rho = rho * np.exp(d - 0.5 * sigma * sigma)
    z = np.zeros_like(rho)
    U = prim_to_cons(rho, z, z, z, p, gamma)
    dx = 1.0 / n
    core = np.broadcast_to(r < r_cells * dx, rho.shape)
    ncore = int(core.sum())
    U[..., 4] += np.where(core, e_factor * abs(W) / (ncore * dx ** 3), 0.0)
    return U


def blobs(n=128, seed=0, turb_sigma=0.3, floor=1e-3):
    """Density field of 1-8 Gaussian or n=1 polytrope blobs (masses
    log-uniform in [0.1, 10], widths in [0.03, 0.15], centres in [0.2, 0.8]^3)
    on a lognormal turbulent background."""
    rng = np.random.default_rng(seed)
    X, Y, Z = centres((n, n, n))
    rho = np.full((n, n, n), floor)
    for _ in range(int(rng.integers(1, 9))):
        m = 10 ** rng.uniform(-1.0, 1.0)
        wdt = rng.uniform(0.03, 0.15)
        c = rng.uniform(0.2, 0.8, 3)
        r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
        if rng.random() < 0.5:
            prof = np.exp(-0.5 * (r / wdt) ** 2)
            norm = (2 * np.pi) ** 1.5 * wdt ** 3
        else:
            Rb = 2.0 * wdt
            x = np.pi * r / Rb
            with np.errstate(invalid="ignore", divide="ignore"):
                prof = np.where(r < Rb, np.where(x > 0, np.sin(x) / np.where(x > 0, x, 1), 1), 0)
            norm = 4.0 * Rb ** 3 / np.pi
        rho = rho + m * prof / norm
    if turb_sigma > 0:
        d = gaussian_field(n, -2.0, turb_sigma, int(rng.integers(1 << 30)))
        rho = rho * np.exp(d - 0.5 * turb_sigma ** 2)
    return rho


def write_profile_file(path, nidx=1.5, R=1.0, rho_c=1.0, G=1.0, gamma=5.0 / 3.0, points=1000):
    """A polytropic star as the radial profile file the reference's stellar
    set-ups read (hydro_solver_3d.c:825-878): '#' header lines, then
    "r rho u" per line, u the specific internal energy p / ((gamma-1) rho),
    from the centre to the surface (Lane-Emden of index nidx, radius R)."""
    xs, th, xi1, _ = lane_emden(nidx)
    alpha = R / xi1
    K = 4.0 * np.pi * G * alpha * alpha * rho_c ** (1.0 - 1.0 / nidx) / (nidx + 1.0)
    pick = np.unique(np.linspace(0, len(xs) - 1, points).astype(int))
    r = xs[pick] * alpha
    rho = rho_c * np.maximum(th[pick], 0.0) ** nidx
    p = K * rho ** (1.0 + 1.0 / nidx)
    u = np.where(rho > 0.0, p / ((gamma - 1.0) * np.where(rho > 0.0, rho, 1.0)), 0.0)
    with open(path, "w") as f:
        f.write(f"# polytrope n = {nidx}, R = {R}, rho_c = {rho_c}, G = {G}, gamma = {gamma}\n")
        f.write("# Columns: radius, density, internal_energy\n")
        for a, b, c in zip(r, rho, u):
            f.write(f"{a:.10e}  {b:.10e}  {c:.10e}\n")
    return path


def write_ic_file(path, U):
    """Binary initial state read by set_initial_conditions_3d_from_deck when
    the deck names it as ic_file: 8-byte magic, int32 nx ny nz, then the
    AoS doubles."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    nx, ny, nz, five = U.shape
    assert five == 5
    with open(path, "wb") as f:
        f.write(IC_MAGIC)
        f.write(np.array([nx, ny, nz], dtype=np.int32).tobytes())
        f.write(U.tobytes())


def read_ic_file(path):
    with open(path, "rb") as f:
        if f.read(8) != IC_MAGIC:
            raise ValueError(f"{path} is not an initial-state file")
        nx, ny, nz = np.frombuffer(f.read(12), dtype=np.int32)
        U = np.frombuffer(f.read(), dtype=np.float64)
    return U.reshape(nx, ny, nz, 5)
