"""The per-pair error band of the BF16 distance tiles, checked on the CPU in
exact arithmetic.

Each BF16 row b carries rho = 1.0079 (sqrt(1.001 sum (b_i - s_i)^2) [+ 2^-23
for Chi2's square-root rows]), s the FP32 row the BF16 values round.  The
tiles treat |b_x . b_y - x . y| <= rho_x + rho_y (+ the FP32 accumulation
term) as certain, where x, y are the exact rows the decision is made on: the
transformed rows (Cosine) or the exact square roots of the transformed rows
(the Chi2 lower bound).  This file restates rho in numpy and checks that
bound on data, on pairs built to make it tight, and against the uniform band
it replaces."""
import numpy as np

from oracle import oracle as O
from sdbscan_amd import datagen

U = 2.0 ** -23


def bf16(a):
    r = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return (((r + 0x7FFF + ((r >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)


def rho(S, sqrt_rows):
    """Per-row bound of the device (FP32 sums of the exact differences)."""
    B = bf16(S)
    df = (B.astype(np.float32) - S.astype(np.float32)).astype(np.float32)
    sq = (df.astype(np.float64) ** 2).sum(1)
    e = np.sqrt(sq * 1.001) + (U if sqrt_rows else 0.0)
    return e * 1.0079, B


def test_cosine_rows_bound_holds_and_is_tighter_than_the_uniform_band():
    X, _ = datagen.mixture(1500, 784, sigma=0.2, seed=5)
    Xt, _ = O.transform(X, 0)
    r, B = rho(Xt, False)
    X64, B64 = Xt.astype(np.float64), B.astype(np.float64)
    err = np.abs(B64 @ B64.T - X64 @ X64.T)
    bound = r[:, None] + r[None, :]
    assert np.all(err <= bound)
    uniform = 2.0 * (2 * 2.0 ** -8 + 2.0 ** -16)  # rounding part of the uniform band
    assert bound.max() < uniform / 3


def test_chi2_sqrt_rows_bound_holds():
    X, _ = datagen.dirichlet(1200, 784, seed=6)
    Xt, _ = O.transform(X, 3)
    S = np.sqrt(Xt)                       # correctly rounded FP32 square roots
    r, B = rho(S, True)
    R64 = np.sqrt(Xt.astype(np.float64))  # the exact square roots
    B64 = B.astype(np.float64)
    err = np.abs(B64 @ B64.T - R64 @ R64.T)
    assert np.all(err <= r[:, None] + r[None, :])


def test_bound_is_tight_for_aligned_rounding_errors():
    """x parallel to the rounding error of y: |x . (b_y - y)| = ||b_y - y||,
    which rho_y must cover (a bound on ||b - s|| alone, or one that drops the
    factor for ||b||, would fail here)."""
    rng = np.random.default_rng(9)
    worst = 0.0
    for t in range(200):
        y = rng.random(784).astype(np.float32)
        y /= np.linalg.norm(y.astype(np.float64)).astype(np.float32)
        dy = bf16(y).astype(np.float64) - y.astype(np.float64)
        x = (dy / np.linalg.norm(dy)).astype(np.float32)
        (rx, ry), Bxy = rho(np.stack([x, y]), False)
        bx, by = Bxy.astype(np.float64)
        err = abs(bx @ by - x.astype(np.float64) @ y.astype(np.float64))
        assert err <= rx + ry
        worst = max(worst, err / (rx + ry))
    assert worst > 0.4


def test_bound_catches_a_dropped_term():
    """Rows whose FP32 square roots are exactly BF16 values, b - s = 0, while
    the exact square roots are not: only the 2^-23 term of the square-root
    rows covers ||b - sqrt(x)||; without it the bound is 0."""
    rng = np.random.default_rng(3)
    fails = 0
    for t in range(100):
        b = bf16(rng.random(784).astype(np.float32) / np.sqrt(784 / 3.0)).astype(np.float64)
        # b^2 is exact in FP32 (8-bit significands); one ulp away from it the
        # exact square root moves by about 2^-25 relative, and for one of the
        # two neighbours sqrtf still returns b
        b2 = (b * b).astype(np.float32)
        up = np.nextafter(b2, np.float32(np.inf))
        dn = np.nextafter(b2, np.float32(0))
        b32 = b.astype(np.float32)
        x = np.where(np.sqrt(up) == b32, up, np.where(np.sqrt(dn) == b32, dn, b2))
        S = np.sqrt(x)
        good, B = rho(S[None, :], True)
        assert np.array_equal(B[0], S) and abs(x.astype(np.float64).sum() - 1.0) < 0.3
        dev = np.linalg.norm(B.astype(np.float64)[0] - np.sqrt(x.astype(np.float64)))
        assert dev <= good[0]
        bare, _ = rho(S[None, :], False)
        fails += dev > bare[0]
    assert fails == 100


def test_entry_to_point_division_magic():
    """a = ent / S by the 40-bit magic ceil(2^40 / S) used on the device:
    exact for every S up to 128 and entries up to 2^31 (the full product's
    bits 40 and up; its low 64 bits alone overflow from S * 2^24 on)."""
    rng = np.random.default_rng(0)
    for S in list(range(2, 129)):
        magic = ((1 << 40) + S - 1) // S
        ents = [int(x) for x in rng.integers(0, 1 << 31, 2000)] + [(1 << 31) - 1, S * (1 << 24), S * (1 << 24) - 1]
        ents += [S * a + s for a in (0, 1, (1 << 31) // S - 1) for s in range(S) if S * a + s < (1 << 31)]
        for e in ents:
            assert (e * magic) >> 40 == e // S, (S, e)
    S = 20
    e = S * (1 << 24) + 3
    assert ((e * (((1 << 40) + S - 1) // S)) & ((1 << 64) - 1)) >> 40 != e // S   # 64-bit product wraps
