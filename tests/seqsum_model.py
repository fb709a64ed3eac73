"""Numpy statement of the device's peak-chunk sequential sum (the method of
hydrocode_astro1_amd/csrc/seqsum.hip), checked on the CPU against a plain
in-order loop.

s = ((0 + x[0]) + x[1]) + ..., every addition rounded.  Write a double s of
binade e (sign and exponent field) as s = I u with I a signed integer,
2^52 <= |I| < 2^53, u = ulp(s).  One addition rounds s + x to the grid of
the result's binade, 2^m u; as long as the result stays in that binade,
adding an even multiple of 2^m u to s shifts the result by exactly that
much (ties to even included), so RN(s + x) - s depends on s only through
I mod 2^(m+1).  Over a run of additions that follows one sequence of
binades (a path), the increment therefore depends only on I mod 2^P, with P
one more than the highest binade of the path above the start binade (0 when
the path never rises above it).

The method cuts the terms into chunks that start where the running sum is
at a peak of its magnitude, so the path rarely rises above the start, and:
  1. block sums and their exclusive prefixes (double-double on the device,
     1024-term blocks) predict the running sum at every term;
  2. in each 4096-term block the predicted peak (largest magnitude) starts
     a chunk, which runs to the next block's peak; P = 1 + how far the next
     peak's binade lies above this one's;
  3. each chunk with P <= 6 is run from 2^P start values near its predicted
     start, one per residue r = I mod 2^P, giving end values E[r];
  4. a serial stitch moves the exact running sum from chunk to chunk:
     r = I mod 2^P of the exact start, o = start - (the start value of
     variant r), end = E[r] + o (both subtractions are exact when start and
     end lie in the predicted binades, which is checked); a chunk whose
     checks fail is added term by term, and a chunk whose terms are all
     zero (both blocks it touches) leaves the sum as it is unless it is NaN
     (the sum starts at +0 and is never -0);
  5. every chunk is run once more from its exact start, in parallel, and
     must reach the next chunk's start; otherwise the whole sum is redone by
     the plain chain.
Step 5 makes the result exact whatever the predictions were: a wrong guess
costs time, never bits.
"""
import math

import numpy as np

B = 4096                 # terms per chunk block (four wavefronts of 64 lanes x 16)
BLOCK2 = 16 * 1024       # the term arrays are padded to a multiple of this
PMAX = 6                 # residue bits tabulated per chunk (2^6 = 64 lanes)


def plain_sum(x):
    s = 0.0
    for v in np.asarray(x, dtype=np.float64).tolist():
        s = s + v
    return s


def _bits(d):
    return int(np.float64(d).view(np.int64)) & 0xFFFFFFFFFFFFFFFF


def _key(d):
    """Sign and exponent field: equal keys = same binade and sign."""
    return _bits(d) >> 52


def _int_mantissa(d):
    b = _bits(d)
    m = (b & ((1 << 52) - 1)) | (1 << 52)
    return -m if b >> 63 else m


def _with_mantissa(i, like):
    b = _bits(like)
    r = (b & (0x7FF << 52)) | ((abs(i) & ((1 << 52) - 1))) | ((1 << 63) if i < 0 else 0)
    return float(np.int64(r - (1 << 64) if r >= 1 << 63 else r).view(np.float64))


def _normal(d):
    e = (_bits(d) >> 52) & 0x7FF
    return 0 < e < 0x7FF


def _run(s, x):
    for v in x:
        s = s + v
    return s


def variant_start(V, P, r):
    """The start value with residue r mod 2^P nearest below V's integer,
    kept inside V's binade."""
    mod = 1 << P
    iv = _int_mantissa(V)
    i = iv - (iv % mod) + r
    if abs(i) < 1 << 52:
        i += mod if i > 0 else -mod
    elif abs(i) >= 1 << 53:
        i -= mod if i > 0 else -mod
    return _with_mantissa(i, V)


def sequential_sum(x, stats=None, guess_noise=None):
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    npad = max(BLOCK2, -(-n // BLOCK2) * BLOCK2)
    t = np.zeros(npad)
    t[:n] = x
    tl = t.tolist()
    nb = npad // B
    with np.errstate(all="ignore"):
        # 1. predicted running sums (exact prefix, rounded once)
        bsum = [math.fsum(tl[j * B:(j + 1) * B]) if np.all(np.isfinite(t[j * B:(j + 1) * B]))
                else float(np.sum(t[j * B:(j + 1) * B])) for j in range(nb)]
        pre = [0.0] * nb
        acc = []
        for j in range(1, nb):
            acc.append(bsum[j - 1])
            pre[j] = math.fsum(acc) if all(map(math.isfinite, acc)) else float(np.sum(acc))
        # 2. peaks
        peak, V = [0] * nb, [0.0] * nb
        for j in range(nb):
            blk = t[j * B:(j + 1) * B]
            v = pre[j] + np.cumsum(blk)  # a prediction only
            a = np.abs(v)
            k = int(np.nanargmax(a)) if np.any(np.isfinite(a)) else 0
            peak[j], V[j] = k, float(v[k])
        if guess_noise is not None:  # predictions spoiled on purpose (tests)
            V = [float(v) for v in np.asarray(V) * guess_noise(nb)]
    zero = [not np.any(t[j * B:(j + 1) * B]) for j in range(nb)]
    zchunk = [zero[j] and (j + 1 == nb or zero[j + 1]) for j in range(nb)]
    first = [j * B + peak[j] + 1 for j in range(nb)]
    last = [first[j + 1] - 1 if j + 1 < nb else npad - 1 for j in range(nb)]
    P = [0] * nb
    for j in range(nb):
        e0 = (_bits(V[j]) >> 52) & 0x7FF
        e1 = (_bits(V[j + 1]) >> 52) & 0x7FF if j + 1 < nb else e0
        ok = _normal(V[j]) and (j + 1 == nb or _normal(V[j + 1]))
        P[j] = max(0, e1 - e0) + 1 if ok else 99
    # 3. tabulated end values
    E = {}
    for j in range(nb):
        if P[j] <= PMAX:
            seg = tl[first[j]:last[j] + 1]
            E[j] = [_run(variant_start(V[j], P[j], r), seg) for r in range(1 << P[j])]
    # 4. stitch
    s = _run(0.0, tl[:first[0]])
    starts = []
    fallback = passed = 0
    for j in range(nb):
        starts.append(s)
        done = False
        if zchunk[j] and not math.isnan(s):
            passed += 1
            done = True
        elif j in E and _key(s) == _key(V[j]):
            r = _int_mantissa(s) % (1 << P[j])
            o = s - variant_start(V[j], P[j], r)
            e = E[j][r]
            end = e + o
            if _normal(e) and _key(end) == _key(e):
                s, done = end, True
        if not done:
            fallback += 1
            s = _run(s, tl[first[j]:last[j] + 1])
    starts.append(s)
    # 5. verification from the exact starts
    bad = sum(1 for j in range(nb)
              if _bits(_run(starts[j], tl[first[j]:last[j] + 1])) != _bits(starts[j + 1]))
    if stats is not None:
        stats.update(chunks=nb, p1=P.count(1), p2=P.count(2), p3=P.count(3),
                     fallback=fallback, verify_fail=bad, zero=passed)
    return plain_sum(tl) if bad else s


def hard_cases(n=200_000, seed=1):
    """Sequences that stress the method: growth, random walks across zero,
    exact halves, cancelling pairs, huge dynamic range, subnormals, and
    column-like swings (terms odd about the middle of each run of 256, as
    the odd moment terms of a centred star are)."""
    rng = np.random.default_rng(seed)
    col = np.tile(np.linspace(-1.0, 1.0, 256), n // 256 + 1)[:n]
    return {
        "positive uniform": rng.random(n),
        "mixed normal": rng.standard_normal(n),
        "lognormal wide range": rng.standard_normal(n) * np.exp(rng.standard_normal(n) * 10),
        "halves (ties)": rng.integers(-5, 6, n).astype(float) * 0.5,
        "big halves": (rng.integers(0, 3, n) + 0.5) * 2.0 ** 50,
        "cancelling pairs": np.repeat(rng.standard_normal(n // 2), 2) * np.tile([1.0, -1.0], n // 2),
        "tiny after huge": np.concatenate([[1e300], rng.standard_normal(n) * 1e-300, [-1e300],
                                           rng.standard_normal(1000)]),
        "subnormals": rng.standard_normal(n) * 1e-310,
        "zeros then terms": np.concatenate([np.zeros(40000), rng.random(n)]),
        "signed zeros": np.where(rng.random(n) < 0.5, -0.0, 0.0),
        "zeros between terms": np.concatenate([rng.random(30000), np.zeros(60000),
                                               -rng.random(30000), np.full(30000, -0.0)]),
        "single": np.ones(1),
        "empty": np.zeros(0),
        "powers of two": np.ldexp(1.0, rng.integers(-60, 60, n)) * rng.choice([-1.0, 1.0], n),
        "ragged length": rng.standard_normal(16384 * 3 + 77) * 1e-3 + 1e-4,
        "column swings": col * (1.0 + 0.01 * rng.standard_normal(n)) * 1e-13,
    }
