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
  3. each chunk with P <= 6 whose start values base + r (r = 0 .. 2^P - 1,
     base = I_V rounded down to a multiple of 2^P) lie in V's binade is run
     from each of them, giving end values E[r] and certificates: how far
     the partial sums stayed from the edges of their binades bounds the
     offset delta by which a start may differ and still end exactly delta
     ulp(start) from E[r] (certificate());
  4. a stitch moves the exact running sum from chunk to chunk: with
     r = I mod 2^P of the exact start and delta = I - (base + r), the end is
     E[r] + delta ulp(start), computed in integers in E[r]'s binade, when
     the start lies in the predicted binade and delta is 0 or within the
     certificate; a chunk where it is not is added term by term, and a
     chunk whose terms are all zero (both blocks it touches) leaves the sum
     as it is unless it is NaN (the sum starts at +0 and is never -0).  The
     chunks are taken 16 at a time: every start residue r of a group's
     first chunk is carried through the group (4a), r mod 2^6 from a base
     near the first chunk's predicted start, which gives a map: a start
     base + r + delta (delta a multiple of 2^6) ends at end_r + delta 2^K_r
     as long as delta 2^K stays a multiple of each later chunk's 2^P (same
     residues; K falls by one for each binade the sum rises) and the offset
     from each chunk's tabulated run within its certificate, which bounds
     delta to a range.  The serial walk then goes from group to group
     through the maps (4b), chunk by chunk where a start is out of its map's
     range, and each chunk's exact start follows from its group's start (4c);
  5. every chunk is run once more from its exact start and must reach the
     next chunk's start (the device does this only when asked, as a check
     of the certificates); otherwise the whole sum is redone by the plain
     chain.
The certificates make step 4 exact by themselves; step 5 checks that.
"""
import math

import numpy as np

B = 4096                 # terms per chunk block (four wavefronts of 64 lanes x 16)
BLOCK2 = 16 * 1024       # the term arrays are padded to a multiple of this
PMAX = 6                 # residue bits tabulated per chunk (2^6 = 64 lanes)
GROUP = 16               # chunks per group of the two-level stitch
PG = 6                   # residue bits of a group's start (one per lane)


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


def _run_cert(s, x):
    """_run, also keeping the smallest distance mm of a partial sum to the
    edges of its binade [2^e, 2^(e+1)) ([0, 2^-1022) for subnormals and
    zeros) and the highest exponent field em met (2047 for inf and NaN)."""
    mm, em = math.inf, 0
    for v in x:
        s = s + v
        if not math.isfinite(s):
            em = 2047
            continue
        a = abs(s)
        if a < 2.0 ** -1022:
            lo, up, f = 0.0, 2.0 ** -1022, 0
        else:
            ex = math.frexp(a)[1]
            lo, up, f = math.ldexp(1.0, ex - 1), math.ldexp(1.0, ex), ex - 1 + 1023
        em = max(em, f)
        mm = min(mm, a - lo, up - a)
    return s, mm, em


def certificate(mm, em, e0, P):
    """The largest |delta| (a multiple of 2^P) for which a run from the start
    v0 + delta ulp(v0) (v0 of exponent field e0) ends exactly delta ulp(v0)
    from the run from v0 whose partial sums kept the distance mm from their
    binades' edges and reached exponent field em at most; -1: only delta =
    0.  Each partial sum of field e <= e0 + P - 1 has spacing g <= 2^(P-1)
    ulp(v0), which D = delta ulp(v0) divides an even number of times, and
    rounding within a binade commutes with such shifts while the shifted
    exact sum stays in the binade: |D| <= m - g for every partial sum."""
    if em > e0 + P - 1:
        return -1
    t = math.ldexp(mm, 1075 - e0) if mm != math.inf else 2.0 ** 58
    allow = (math.floor(t) if t >= 1.0 else 0) - (1 << (P - 1))
    return min(allow, 1 << 58) if allow >= 0 else -1


def _chunk_end(I, es, P, base, e, allow):
    """The end of a chunk from the start integer I (exponent field es, in the
    chunk's predicted binade) given e, the end tabulated from base + r (r = I
    mod 2^P), and that run's certificate: e's integer plus delta = I - base -
    r scaled to e's binade, or None when delta is neither 0 nor within the
    certificate, e is not normal, the scaling is not exact or could overflow,
    or the end leaves e's binade.  Returns (end, k), k = es - e's exponent
    field."""
    r = I % (1 << P)
    delta = I - base - r
    if delta != 0 and abs(delta) > allow:
        return None
    ee = (_bits(e) >> 52) & 0x7FF
    if ee == 0 or ee == 0x7FF:
        return None
    k = es - ee
    if k >= 0:
        if k > 10 or abs(delta) >= 1 << 50:
            return None
        D = delta << k
    else:
        if -k >= P:
            return None
        D = delta >> -k
    IE = _int_mantissa(e)
    Ie = IE + D
    if (Ie < 0) != (IE < 0) or not (1 << 52) <= abs(Ie) < (1 << 53):
        return None
    return _with_mantissa(Ie, e), k


def sequential_sum(x, stats=None, guess_noise=None, group=GROUP):
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
    # 3. tabulated end values, and what the stitch needs of each chunk: the
    #    start values' common base (their residues are 0 .. 2^P - 1) and
    #    whether all of them lie in V's binade
    #    whether all of them lie in V's binade; with each end value its
    #    certificate
    es = [(_bits(v) >> 52) & 0x7FF for v in V]
    E, Z, base, okc = {}, {}, [0] * nb, [False] * nb
    for j in range(nb):
        if P[j] <= PMAX:
            iv = _int_mantissa(V[j])
            base[j] = iv - iv % (1 << P[j])
            okc[j] = all((1 << 52) <= abs(i) < (1 << 53) for i in (base[j], base[j] + (1 << P[j]) - 1))
        if okc[j] and not zchunk[j]:
            seg = tl[first[j]:last[j] + 1]
            runs = [_run_cert(_with_mantissa(base[j] + r, V[j]), seg) for r in range(1 << P[j])]
            E[j] = [e for e, _, _ in runs]
            Z[j] = [certificate(mm, em, es[j], P[j]) for _, mm, em in runs]
    cnt = dict(fallback=0, zero=0)

    def step(s, j):
        """One chunk of the chunk-by-chunk stitch from the exact start s."""
        if zchunk[j] and not math.isnan(s):
            cnt["zero"] += 1
            return s
        if okc[j] and _key(s) == _key(V[j]):
            I = _int_mantissa(s)
            r = I % (1 << P[j])
            res = _chunk_end(I, es[j], P[j], base[j], E[j][r], Z[j][r])
            if res is not None:
                return res[0]
        cnt["fallback"] += 1
        return _run(s, tl[first[j]:last[j] + 1])

    # 4a. group maps: for each group of chunks and start residue r mod 2^PG
    #     (from a base near its first chunk's predicted start), the end from
    #     base + r, the shift K and the range [lo, hi]: from base + r + delta
    #     (delta a multiple of 2^PG in [lo, hi]) the group ends at
    #     end + delta 2^K; the range keeps the offset from each chunk's
    #     tabulated run, d + delta 2^K, within that run's certificate
    ng = -(-nb // group)
    maps, allzero, gbase = [], [], []
    for g in range(ng):
        cb, ce = g * group, min(nb, (g + 1) * group)
        allzero.append(all(zchunk[cb:ce]))
        row = {}
        gb = None
        if _normal(V[cb]):
            iv = _int_mantissa(V[cb])
            gb = iv - iv % (1 << PG)
            if not all((1 << 52) <= abs(i) < (1 << 53) for i in (gb, gb + (1 << PG) - 1)):
                gb = None
        gbase.append(gb)
        if not allzero[-1] and gb is not None:
            for r in range(1 << PG):
                v, K, lo, hi, valid = _with_mantissa(gb + r, V[cb]), 0, -(1 << 62), 1 << 62, True
                for j in range(cb, ce):
                    if zchunk[j]:
                        continue
                    if not okc[j] or _key(v) != _key(V[j]) or PG + K < P[j]:
                        valid = False
                        break
                    I = _int_mantissa(v)
                    rr = I % (1 << P[j])
                    res = _chunk_end(I, es[j], P[j], base[j], E[j][rr], Z[j][rr])
                    if res is None:
                        valid = False
                        break
                    # the start's offset from this run, d + delta 2^K, within
                    # [-A, A]: delta in [(-A - d) / 2^K, (A - d) / 2^K]
                    A, d = max(Z[j][rr], 0), I - base[j] - rr
                    a, b = -A - d, A - d
                    if K >= 0:
                        lo, hi = max(lo, -((-a) >> K)), min(hi, b >> K)
                    else:
                        lo, hi = max(lo, a << -K), min(hi, b << -K)
                    v, k = res
                    K += k
                if valid and lo <= hi:
                    row[r] = (v, K, lo, hi)
        maps.append(row)
    # 4b. the exact sum from group to group: through the map where the start
    #     lies in the first chunk's binade within the room, chunk by chunk
    #     otherwise
    s = _run(0.0, tl[:first[0]])
    gstart, walked = [], 0
    for g in range(ng):
        gstart.append(s)
        cb, ce = g * group, min(nb, (g + 1) * group)
        if allzero[g] and not math.isnan(s):
            continue
        done = False
        if gbase[g] is not None and _key(s) == _key(V[cb]):
            I = _int_mantissa(s)
            r = I % (1 << PG)
            delta = I - gbase[g] - r
            if r in maps[g] and maps[g][r][2] <= delta <= maps[g][r][3]:
                v, K, _, _ = maps[g][r]
                s = _with_mantissa(_int_mantissa(v) + (delta << K if K >= 0 else delta >> -K), v)
                done = True
        if not done:
            walked += 1
            for j in range(cb, ce):
                s = step(s, j)
    total = s
    # 4c. every chunk's exact start, chunk by chunk from its group's start
    cnt.update(fallback=0, zero=0)
    starts = []
    for g in range(ng):
        s = gstart[g]
        for j in range(g * group, min(nb, (g + 1) * group)):
            starts.append(s)
            s = step(s, j)
    starts.append(total)
    # 5. verification from the exact starts
    bad = sum(1 for j in range(nb)
              if _bits(_run(starts[j], tl[first[j]:last[j] + 1])) != _bits(starts[j + 1]))
    if stats is not None:
        stats.update(chunks=nb, p1=P.count(1), p2=P.count(2), p3=P.count(3),
                     fallback=cnt["fallback"], verify_fail=bad, zero=cnt["zero"], groups=ng,
                     walked=walked)
    return plain_sum(tl) if bad else total


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
