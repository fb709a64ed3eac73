"""Numpy statement of the device's block-composite sequential sum (the
composite method of hydrocode_astro1_amd/csrc/seqsum.hip), checked on the
CPU against a plain loop (np.cumsum adds in order, rounding every step).

s = ((0 + x[0]) + x[1]) + ...: while the running sum stays in one binade
[2^k, 2^(k+1)) of one sign, every addition rounds to the grid u = 2^(k-52),
so with S = s/u an integer, s + x rounds to u (S + rint(x/u)) exactly when
the exact S + x/u stays in [2^52, 2^53), i.e. S + floor(x/u) does (halves
are left to the chain: their rounding depends on the parity of S).  A block
is summarised by the increment D = sum rint(x/u), and lo / hi, the extremes
of (increments so far + floor(x_j/u)); it applies to an entry S when
S + lo >= 2^52 and S + hi < 2^53.  Binades are guessed from approximate
prefix sums; a guess that does not hold only sends the walk down a level,
to chunks added term by term.
"""
import numpy as np

CHUNK = 16              # terms per chunk (a lane on the device)
LANES = 64              # chunks per 1024-block (a wavefront)
B1 = CHUNK * LANES      # 1024
B2 = 16 * B1            # 16384
TWO52, TWO53 = 2 ** 52, 2 ** 53
LIM = 2 ** 55           # |D|, lo, hi past this can never apply


def _keys(P):
    _, e = np.frexp(P)
    k = e - 1
    ok = np.isfinite(P) & (P != 0) & (k >= -960) & (k <= 1000)
    return np.where(P < 0, -1, 1), np.where(ok, k, 0), ok


def _chunks(C, sig, k, ok):
    """Per chunk D, lo, hi, valid, zero for chunks C [nchunk][CHUNK]."""
    with np.errstate(over="ignore", invalid="ignore"):
        sx = sig[:, None] * C
        t = np.ldexp(sx, (52 - k)[:, None])
        at = np.abs(t)
        small = at < 0.5
        fl = np.floor(t)
        bad = ~(at < 2.0 ** 54) | (~small & (t - fl == 0.5))
        f = np.where(small, np.where(sx < 0, -1.0, 0.0), fl)
        d = np.where(small, 0.0, np.rint(t))
    f = np.where(bad, 0, f).astype(np.int64)
    d = np.where(bad, 0, d).astype(np.int64)
    E = np.concatenate([np.zeros((len(C), 1), np.int64), np.cumsum(d, axis=1)[:, :-1]], axis=1)
    D, lo, hi = d.sum(1), (E + f).min(1), (E + f).max(1)
    valid = ok & ~bad.any(1) & (np.abs(D) < LIM) & (lo > -LIM) & (hi < LIM)
    zero = (C == 0).all(1)
    return D, lo, hi, valid, zero


def _compose(D, lo, hi, valid, zero, sig, k, g):
    """Blocks of g consecutive descriptors: valid when every non-zero child
    is valid with one binade and sign."""
    sh = (-1, g)
    D, lo, hi, valid, zero, sig, k = (a.reshape(sh) for a in (D, lo, hi, valid, zero, sig, k))
    Dz = np.where(zero, 0, D)
    E = np.concatenate([np.zeros((len(D), 1), np.int64), np.cumsum(Dz, axis=1)[:, :-1]], axis=1)
    lo2 = np.where(zero, LIM, E + lo).min(1)
    hi2 = np.where(zero, -LIM, E + hi).max(1)
    D2 = Dz.sum(1)
    code = np.where(zero, 0, k * 2 + (sig < 0))
    first = np.argmax(~zero, axis=1)
    c0 = code[np.arange(len(code)), first]
    same = np.where(zero, True, code == c0[:, None]).all(1)
    z2 = zero.all(1)
    v2 = (np.where(zero, True, valid).all(1) & same & (np.abs(D2) < LIM) & (lo2 > -LIM)
          & (hi2 < LIM))
    return D2, lo2, hi2, v2, z2, np.where(c0 & 1, -1, 1), c0 >> 1


def prepare(x, guess_noise=None):
    """Descriptors of the three levels for the zero-padded terms.
    guess_noise (tests only) scales each approximate prefix by a factor, to
    make binade guesses wrong on purpose."""
    n = len(x)
    npad = max(1, -(-n // B2)) * B2
    xp = np.zeros(npad)
    xp[:n] = x
    C = xp.reshape(-1, CHUNK)
    P = np.concatenate([[0.0], np.cumsum(C.sum(1))[:-1]])  # approximate, any order will do
    if guess_noise is not None:
        P = P * guess_noise(len(P))
    sig, k, ok = _keys(P)
    L0 = _chunks(C, sig, k, ok) + (sig, k)
    L1 = _compose(*L0, LANES)
    L2 = _compose(*L1, 16)
    return xp, L0, L1, L2


def apply(desc, i, s):
    """s moved over block i, or None when that is not exact."""
    D, lo, hi, valid, zero, sig, k = desc
    if zero[i]:
        return s
    if not valid[i]:
        return None
    kk = int(k[i])
    a = float(sig[i]) * s
    if not (np.ldexp(1.0, kk) <= a <= np.ldexp(1.0, kk + 1)):
        return None
    S = int(np.ldexp(a, 52 - kk))
    if S + int(lo[i]) < TWO52 or S + int(hi[i]) >= TWO53:
        return None
    return float(sig[i]) * float(np.ldexp(float(S + int(D[i])), kk - 52))


def segments(L0, b1):
    """The runs of chunks of 1024-block b1 that share a binade and sign with
    every non-zero chunk valid (zero chunks join the run they fall in, an
    invalid chunk stands alone): [(start, end, descriptor)]."""
    D, lo, hi, valid, zero, sig, k = L0
    c0 = LANES * b1
    code = k * 2 + (sig < 0)
    starts, prev = [0], -1
    for c in range(c0, c0 + LANES):
        if not zero[c]:
            if c > c0 and prev >= 0 and (not valid[c] or not valid[prev] or code[c] != code[prev]):
                starts.append(c - c0)
            prev = c
    out = []
    for a, b in zip(starts, starts[1:] + [LANES]):
        idx = np.arange(c0 + a, c0 + b)
        one = _compose(*(np.asarray(v)[idx] for v in L0), len(idx))
        out.append((a, b, tuple(np.asarray(v)[:1] for v in one)))
    return out


def sequential_sum(x, stats=None, guess_noise=None):
    """The walk: 16384-blocks, else their 1024-blocks, else the segments of
    the 1024-block (those of three chunks or more), else the segment's terms
    one by one."""
    xp, L0, L1, L2 = prepare(np.asarray(x, dtype=np.float64), guess_noise)
    s = 0.0
    st = dict(l2=0, l1=0, seg=0, chain=0)
    for b2 in range(len(L2[0])):
        r = apply(L2, b2, s)
        if r is not None:
            s, st["l2"] = r, st["l2"] + 1
            continue
        for b1 in range(16 * b2, 16 * b2 + 16):
            r = apply(L1, b1, s)
            if r is not None:
                s, st["l1"] = r, st["l1"] + 1
                continue
            for a, b, sd in segments(L0, b1):
                r = apply(sd, 0, s) if b - a >= 3 else None
                if r is not None:
                    s, st["seg"] = r, st["seg"] + 1
                    continue
                st["chain"] += b - a
                for e in range(CHUNK * (LANES * b1 + a), CHUNK * (LANES * b1 + b)):
                    s = s + xp[e]
    if stats is not None:
        stats.update(st)
    return s


def plain_sum(x):
    """The loop itself: np.cumsum rounds after every addition, in order."""
    x = np.asarray(x, dtype=np.float64)
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def hard_cases(n=200_000, seed=1):
    """Sequences that stress the method: growth, random walks across zero,
    exact halves, cancelling pairs, huge dynamic range, subnormals."""
    rng = np.random.default_rng(seed)
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
        "single": np.ones(1),
        "empty": np.zeros(0),
        "powers of two": np.ldexp(1.0, rng.integers(-60, 60, n)) * rng.choice([-1.0, 1.0], n),
        "ragged length": rng.standard_normal(16384 * 3 + 77) * 1e-3 + 1e-4,
    }
