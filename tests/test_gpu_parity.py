"""Parity of the HIP path (through the C ABI) with the oracle on the same inputs.

Each stage is fed the same inputs on both sides (staged parity); the bar is
bit-exact for projections' derived index sets, core flags, degrees and labels.
"""
import numpy as np
import pytest

import sdbscan_amd as S
from oracle import oracle as O
from sdbscan_amd import datagen

pytestmark = pytest.mark.gpu


def run_index(ctx, X, metric=S.COSINE, n_proj=1024, K=5, M=50, seed=42, ker=None, sigma=1.0,
              L=0.4, precision=S.FP32):
    n, d = X.shape
    p = ctx.set_params(n, d, n_proj=n_proj, topK=K, topM=M, metric=metric, ker_n_features=ker,
                       ker_sigma=sigma, ker_intervalSampling=L, precision=precision)
    bits = O.hd3_bits(3 * p.fhtDim, seed)
    R = None
    if metric in (S.L1, S.L2):
        half = p.ker_n_features // 2
        R = (O.gauss if metric == S.L2 else O.cauchy)(half, d, 0.0, 1.0 / sigma, seed)
    ctx.load(X)
    ctx.build_index(bits, R)
    return p, bits, R


def oracle_index(Xt, metric, p, bits, R):
    E = O.embed(Xt, metric, p.ker_n_features, p.ker_intervalSampling, R)
    proj = O.project(E, bits, p.fhtDim, p.n_proj)
    return proj, O.topk(proj, p.topK), O.topm(proj, p.topM)


def sets_equal(a, b):
    return all(set(x) == set(y) for x, y in zip(a, b))


def index_equal(tk, tko, tm, tmo, K, M):
    """Index rows equal as sets (close and far halves).  The order among
    equal projection values is the heap's internal order in the reference
    (unspecified), so only the sets are compared."""
    return (sets_equal(tk[:, :K], tko[:, :K]) and sets_equal(tk[:, K:], tko[:, K:])
            and sets_equal(tm[:, :M], tmo[:, :M]) and sets_equal(tm[:, M:], tmo[:, M:]))


@pytest.mark.parametrize("metric,d,ker", [
    (S.COSINE, 784, None),   # fused into the index passes
    (S.CHI2, 784, 1024),     # fused (square-root features)
    (S.CHI2, 100, 512),      # transform kernel, then the LDS embedding
])
def test_transform_bit_exact(ctx, metric, d, ker):
    if metric == S.COSINE:
        X, _ = datagen.mixture(3000, d, seed=7)
    else:
        X, _ = datagen.dirichlet(3000, d, seed=7)
    X = X * np.float32(3.7)
    print(f"transform metric {metric} n=3000 d={d} ker={ker}", flush=True)
    run_index(ctx, X, metric=metric, ker=ker, n_proj=256)
    Xg = ctx.get_transformed()
    Xo, rc = O.transform(X, metric)
    assert rc == 0
    assert np.array_equal(Xg.view(np.uint32), Xo.view(np.uint32))


def test_bf16_rows_round_to_nearest_even(ctx):
    """The BF16 copy of the transformed rows is the round-to-nearest-even of
    the FP32 rows, exact halfway cases included (the recheck band of the BF16
    tiles assumes half-ulp rounding)."""
    X, _ = datagen.mixture(3000, 784, seed=19)
    X = X * np.float32(0.37)
    print("bf16 rows cosine n=3000 d=784", flush=True)
    run_index(ctx, X, precision=S.BF16)
    Xn = ctx.get_transformed()
    Xb = ctx.get_transformed_bf16()
    u = Xn.view(np.uint32).astype(np.uint64)
    low = u & 0xFFFF
    up = (u >> 16) + ((low > 0x8000) | ((low == 0x8000) & ((u >> 16) & 1 == 1)))
    assert np.count_nonzero(low == 0x8000) > 0   # halfway cases present
    assert np.array_equal(Xb, up.astype(np.uint16))


def test_index_matches_oracle(ctx):
    X, _ = datagen.mixture(4000, 784, seed=11)
    print("index cosine n=4000 K=5 M=50", flush=True)
    p, bits, R = run_index(ctx, X, K=5, M=50)
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    st = ctx.stats()
    print("stats", st, flush=True)
    assert index_equal(tk, tko, tm, tmo, p.topK, p.topM)
    # without ties among the selected values the order is the reference's too
    assert np.mean(np.all(tk == tko, axis=1)) > 0.99


@pytest.mark.parametrize("n_proj,d,K,M", [(1000, 100, 7, 40), (200, 33, 3, 25), (1024, 2000, 6, 30)])
def test_index_ragged_projection_counts(ctx, n_proj, d, K, M):
    """n_proj not a multiple of 64 (the last register of a lane half valid),
    n_proj far below fhtDim, and rows longer than fhtDim's 1024 (d = 2000,
    fhtDim 2048): index, cores and labels equal the oracle's."""
    X, _ = datagen.mixture(3000, d, k=5, sigma=0.2, seed=200 + d)
    print(f"ragged index n=3000 d={d} n_proj={n_proj} K={K} M={M}", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, 0.3, 10, K=K, M=M, n_proj=n_proj)
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    assert index_equal(tk, tko, tm, tmo, K, M)


def test_core_and_labels_match_oracle(ctx):
    X, _ = datagen.mixture(5000, 784, seed=13)
    print("core+labels cosine n=5000 eps=0.25 minPts=50", flush=True)
    p, bits, R = run_index(ctx, X, K=5, M=50)
    core, deg = ctx.find_core(0.25, 50)
    lab, nc = ctx.form_clusters()
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    oc, od, (ptr, idx, dd), near, ndist = O.find_core(Xt, S.COSINE, tk, tm, 0.25, 50, mode=1)
    olab, onc = O.form_cluster(oc, ptr, idx)
    print("stats", ctx.stats(), "near", near, flush=True)
    assert np.array_equal(deg, od)
    assert np.array_equal(core, oc)
    assert np.array_equal(lab, olab)
    assert nc == onc


def test_bf16_matches_fp32(ctx):
    X, _ = datagen.mixture(5000, 784, seed=17)
    print("bf16 vs fp32 n=5000", flush=True)
    run_index(ctx, X, precision=S.FP32)
    c32, d32 = ctx.find_core(0.25, 50)
    l32, n32 = ctx.form_clusters()
    run_index(ctx, X, precision=S.BF16)
    c16, d16 = ctx.find_core(0.25, 50)
    l16, n16 = ctx.form_clusters()
    print("bf16 stats", ctx.stats(), flush=True)
    assert np.array_equal(d32, d16)
    assert np.array_equal(c32, c16)
    assert np.array_equal(l32, l16) and n32 == n16


@pytest.mark.parametrize("n,d,K,M,eps", [
    (6000, 784, 10, 100, 0.33),   # 25 k-steps, 7 candidate tiles (the 8M x 784 shape)
    (6000, 784, 5, 50, 0.33),     # 4 candidate tiles, 2 mask words
    (5000, 300, 8, 64, 0.30),     # 10 k-steps padded to 25
    (5000, 128, 8, 50, 0.30),     # 4 k-steps padded to 8
])
def test_bf16_tiles_match_oracle(ctx, n, d, K, M, eps):
    # eps sits inside the bulk of the within-cluster distances, so many pairs
    # fall in the BF16 error band and go through the FP32/FP64 recheck
    X, _ = datagen.mixture(n, d, sigma=0.3, seed=81 + d)
    print(f"bf16 tiles n={n} d={d} K={K} M={M} eps={eps}", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, eps, 10, K=K, M=M, precision=S.BF16)
    print("queue", st["n_queue"], "fp32", st["n_recheck32"], "fp64", st["n_recheck64"], flush=True)
    assert st["n_queue"] > 0


def full_check(ctx, X, metric, eps, minPts, **kw):
    """GPU index + core + labels against the oracle fed the GPU's transformed
    points and index (staged parity); returns the GPU stats."""
    p, bits, R = run_index(ctx, X, metric=metric, **kw)
    core, deg = ctx.find_core(eps, minPts)
    lab, nc = ctx.form_clusters()
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    oc, od, (ptr, idx, dd), near, ndist = O.find_core(Xt, metric, tk, tm, eps, minPts, mode=1)
    olab, onc = O.form_cluster(oc, ptr, idx)
    st = ctx.stats()
    print("stats", {k: st[k] for k in ("n_hubs", "n_queue", "n_recheck64", "n_tie_points",
                                       "n_tie_rows", "n_core", "n_components")}, "near", near, flush=True)
    assert np.array_equal(deg, od)
    assert np.array_equal(core, oc)
    assert np.array_equal(lab, olab) and nc == onc
    return p, bits, R, tk, tm, Xt, st


def test_c0_parameters_k10_m100(ctx):
    X, _ = datagen.mixture(20000, 784, seed=1000)
    print("c0-like n=20000 K=10 M=100 eps=0.25 minPts=50", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, 0.25, 50, K=10, M=100)
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    assert index_equal(tk, tko, tm, tmo, 10, 100)


@pytest.mark.parametrize("precision", [S.FP32, S.BF16])
@pytest.mark.parametrize("case", ["c0", "chi2"])
def test_end_to_end_matches_reference_arithmetic_fit(ctx, case, precision):
    """The whole path on raw rows against the oracle's full fit, which decides
    with the reference's FP32 arithmetic (sequential sums): c0 exactly
    (20,000 x 784 cosine, topK 5, topM 50, eps 0.25, minPts 50), and 10,000
    Chi2 histograms at the 8M x 784 Chi2 parameters with eps 0.47, minPts 10
    (at eps 0.2 every point is noise).  On these inputs no pair's FP32 and
    FP64 decisions differ, so labels, cluster counts and core flags must be
    equal."""
    if case == "c0":
        X, _ = datagen.mixture(20000, 784, seed=1000)
        metric, K, M, ker, eps, mp = S.COSINE, 5, 50, 784, 0.25, 50
    else:
        X, _ = datagen.dirichlet(10000, 784, seed=1004)
        metric, K, M, ker, eps, mp = S.CHI2, 10, 100, 1024, 0.47, 10
    n, d = X.shape
    print(f"end to end {case} n={n} precision={precision}", flush=True)
    p = ctx.set_params(n, d, n_proj=1024, topK=K, topM=M, metric=metric, ker_n_features=ker,
                       precision=precision)
    bits = S.hd3_bits(3 * p.fhtDim, 42)
    ctx.load(X)
    ctx.build_index(bits)
    core, _ = ctx.find_core(eps, mp)
    lab, nc = ctx.form_clusters()
    olab, onc, ocore, _ = O.fit_sdbscan(X, metric, 1024, K, M, ker, 0.4, None, bits, p.fhtDim, eps, mp,
                                        O.max_threads())
    print("clusters", nc, "oracle", onc, "cores", int(core.sum()), flush=True)
    assert np.array_equal(core, ocore)
    assert np.array_equal(lab, olab) and nc == onc


@pytest.mark.parametrize("n,d,K,M,n_proj,precision", [
    (200, 16, 3, 200, 64, S.FP32),   # topM = n: every list holds every point (D = 64)
    (200, 16, 3, 200, 64, S.BF16),
    (64, 40, 1, 64, 128, S.FP32),    # one slot pair per point, two mask words
])
def test_all_points_in_every_list_is_exact_dbscan(ctx, n, d, K, M, n_proj, precision):
    """topM = n makes every point a candidate of every other: the labels are
    exact DBSCAN's (sklearn pins the core points), the pair tables are
    longer than the sorted segments (the scan path), and every point is a
    member of every list."""
    from sklearn.cluster import DBSCAN
    X, _ = datagen.mixture(n, d, k=3, sigma=0.3, seed=401 + n)
    eps, mp = 0.3, 5
    print(f"all points candidates n={n} d={d} K={K} M={M} n_proj={n_proj} precision={precision}", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, eps, mp, K=K, M=M, n_proj=n_proj, precision=precision)
    core, _ = ctx.find_core(eps, mp)
    sk = DBSCAN(eps=eps, min_samples=mp, metric="cosine").fit(Xt.astype(np.float64))
    sk_core = np.zeros(n, np.uint8)
    sk_core[sk.core_sample_indices_] = 1
    assert np.array_equal(core, sk_core)


def test_duplicate_points_replay_tied_rows(ctx):
    X, _ = datagen.mixture(3000, 64, k=4, seed=21)
    X[1000:1500] = X[:500]          # exact duplicates: tied projections in every row
    print("duplicates n=3000 d=64 K=5 M=40", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, 0.2, 10, K=5, M=40, n_proj=256)
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    # rows tied across the top-M boundary are replayed with the reference heap
    assert st["n_tie_rows"] > 0
    assert index_equal(tk, tko, tm, tmo, p.topK, p.topM)


def test_one_hot_points_tied_topk(ctx):
    n, d = 2000, 96
    rng = np.random.default_rng(3)
    X, _ = datagen.mixture(n, d, k=3, seed=23)
    onehot = rng.integers(0, d, 300)
    X[:300] = 0.0
    X[np.arange(300), onehot] = 1.0      # integer-valued projections with many ties
    print("one-hot rows n=2000 d=96 K=8 M=30", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, 0.3, 5, K=8, M=30, n_proj=128)
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    assert st["n_tie_points"] > 0
    assert index_equal(tk, tko, tm, tmo, p.topK, p.topM)


@pytest.mark.parametrize("metric,eps", [(S.L2, 4.0), (S.L1, 25.0)])
def test_l1_l2_random_features(ctx, metric, eps):
    X, _ = datagen.blobs(6000, 51, k=6, seed=31)
    print(f"metric {metric} RFF n=6000 d=51 K=10 M=60", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, metric, eps, 20, K=10, M=60, ker=1024, sigma=4.0)
    # projections: the oracle re-derives them from the same inputs; cos/sin of
    # the device and of libm may differ in the last ulp, so the index is
    # compared as sets with a tolerance on the embedded values
    E = O.embed(Xt, metric, p.ker_n_features, p.ker_intervalSampling, R)
    proj = O.project(E, bits, p.fhtDim, p.n_proj)
    agree = np.mean([set(a) == set(b) for a, b in zip(tk[:, :p.topK], O.topk(proj, p.topK)[:, :p.topK])])
    print("top-K agreement with libm embedding", agree, flush=True)
    assert agree > 0.99


@pytest.mark.parametrize("metric,d,M,eps", [
    (S.COSINE, 64, 200, 0.3),    # Cosine with topM > 128: VALU tiles, 4 candidate columns per lane
    (S.L1, 40, 150, 25.0),       # 3 columns per lane, rows padded to 4
    (S.L2, 300, 120, 10.6),      # rows longer than 256: k-chunked VALU tiles
])
def test_valu_tile_shapes(ctx, metric, d, M, eps):
    if metric == S.COSINE:
        X, _ = datagen.mixture(4000, d, k=5, sigma=0.3, seed=301)
        kw = {}
    else:
        X, _ = datagen.blobs(4000, d, k=5, seed=301)
        kw = dict(ker=1024, sigma=4.0 if metric == S.L1 else 6.0)
    print(f"valu tiles metric {metric} d={d} M={M} eps={eps}", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, metric, eps, 10, K=6, M=M, n_proj=256, **kw)
    print("cores", st["n_core"], "queue", st["n_queue"], flush=True)


@pytest.mark.parametrize("metric,eps", [(S.CHI2, 0.35), (S.JS, 0.3)])
@pytest.mark.parametrize("ker", [512, 256])
def test_chi2_js(ctx, metric, eps, ker):
    # ker 512 >= 3 d: sampled components (LDS embedding after the transform
    # kernel); ker 256 < 3 d: square-root features with the transform fused
    X, _ = datagen.dirichlet(5000, 100, k=5, seed=41)
    print(f"metric {metric} n=5000 d=100 ker={ker}", flush=True)
    full_check(ctx, X, metric, eps, 20, K=8, M=50, ker=ker, n_proj=256)


@pytest.mark.parametrize("n,d,K,M,eps,minPts", [
    (6000, 784, 10, 100, 0.47, 10),   # single-chunk BF16 tiles (the 8M x 784 Chi2 shape)
    (5000, 100, 8, 50, 0.25, 20),     # 4 k-steps padded to 8
    (3000, 1000, 5, 50, 0.5, 5),      # rows longer than one LDS chunk: k-chunked tiles
])
def test_chi2_bf16_prefilter_matches_oracle(ctx, n, d, K, M, eps, minPts):
    """Chi2 with BF16 rows: the MFMA tiles only prune (1 - sum sqrt(x y) bounds
    chi2 from below) and every surviving pair is decided by the FP32/FP64
    recheck, so degrees, cores and labels equal the oracle's and the FP32
    path's."""
    X, _ = datagen.dirichlet(n, d, k=10 if d >= 784 else 5, seed=61 + d)
    print(f"chi2 bf16 prefilter n={n} d={d} K={K} M={M} eps={eps}", flush=True)
    ker = 1024 if d > 512 else 512
    full_check(ctx, X, S.CHI2, eps, minPts, K=K, M=M, ker=ker, n_proj=256, precision=S.FP32)
    c32, d32 = ctx.find_core(eps, minPts)
    l32, n32 = ctx.form_clusters()
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.CHI2, eps, minPts, K=K, M=M, ker=ker, n_proj=256,
                                            precision=S.BF16)
    print("queue", st["n_queue"], "fp32", st["n_recheck32"], "fp64", st["n_recheck64"], "band", st["band"],
          flush=True)
    assert st["n_queue"] > 0 and st["n_recheck32"] == st["n_queue"]
    c16, d16 = ctx.find_core(eps, minPts)
    l16, n16 = ctx.form_clusters()
    assert np.array_equal(d32, d16) and np.array_equal(c32, c16)
    assert np.array_equal(l32, l16) and n32 == n16
    # the BF16 rows are the square roots of the sum-normalised rows, rounded
    # to BF16 (within one BF16 ulp of numpy's correctly rounded result)
    xb = ctx.get_transformed_bf16().astype(np.int32)
    r = np.sqrt(Xt).view(np.uint32).astype(np.uint64)
    rne = ((r + 0x7FFF + ((r >> 16) & 1)) >> 16).astype(np.int32)
    assert np.abs(xb - rne).max() <= 1


@pytest.mark.parametrize("shift", [0.0, 100.0, 3000.0])
def test_l2_mfma_tiles_and_valu_fallback_match_oracle(ctx, shift):
    """L2 distances on the FP32 MFMA tiles through |x|^2 + |y|^2 - 2 x.y with
    a per-pair band of kappa (|x|^2 + |y|^2) (shift 0; shift 100 widens the
    band and sends many pairs to the FP64 recheck), and on the VALU tiles
    when rows far from the origin would widen it past 5% of eps^2 (shift
    3000).  Degrees, cores and labels equal the oracle's FP64 decision in
    every case; the reported band tells the two tile kinds apart."""
    X, _ = datagen.blobs(6000, 51, k=12, noise=0.05, seed=1001)
    sigma = datagen.median_pairwise_l2(X, 4000, seed=1001)
    eps = float(np.float32(datagen.l2_eps_for_embedded_cosine(sigma, 0.3)))
    X = (X + np.float32(shift)).astype(np.float32)
    print(f"L2 tiles shift={shift} n=6000 d=51 eps={eps:.3f}", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.L2, eps, 20, K=10, M=100, ker=1024, sigma=sigma)
    band_valu = 2.0 * eps * (51 + 8) * 2.0 ** -23
    print("band", st["band"], "valu band", band_valu, "queue", st["n_queue"], flush=True)
    # the MFMA tiles report their widest per-pair band, (2 kappa max|x|^2 +
    # eps0) / (2 eps); the VALU tiles report 2 eps (d + 8) u
    if shift < 1000:
        assert abs(st["band"] - band_valu) > 1e-6 * band_valu
    else:
        assert abs(st["band"] - band_valu) <= 1e-12 * band_valu


def test_sorted_recheck_queue_matches_oracle(ctx):
    """A populated error band (mixture noise 0.2, eps 0.27 inside the
    within-cluster distances): more than 2^14 pairs go through the recheck,
    which decides them in candidate order (the queue sorted by candidate
    point); degrees, cores and labels equal the oracle's FP64 decision."""
    X, _ = datagen.mixture(40000, 784, sigma=0.2, seed=1002)
    print("populated band n=40000 cosine bf16 eps 0.27", flush=True)
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.COSINE, 0.27, 20, K=10, M=100, precision=S.BF16)
    print("queue", st["n_queue"], "fp32", st["n_recheck32"], "fp64", st["n_recheck64"], flush=True)
    assert st["n_queue"] >= 1 << 14


def test_chi2_bf16_queue_overflow_falls_back_to_fp32_tiles(ctx, monkeypatch):
    """More pairs near eps than the recheck queue may hold: the pass is
    redone on the FP32 tiles, with the same decisions."""
    X, _ = datagen.dirichlet(5000, 100, k=5, seed=161)
    full_check(ctx, X, S.CHI2, 0.25, 20, K=8, M=50, ker=256, n_proj=256, precision=S.BF16)
    nq = ctx.stats()["n_queue"]
    assert nq > 1000
    monkeypatch.setenv("SDB_QUEUE_MAX", "1000")
    p, bits, R, tk, tm, Xt, st = full_check(ctx, X, S.CHI2, 0.25, 20, K=8, M=50, ker=256, n_proj=256,
                                            precision=S.BF16)
    print("queue limit 1000: queued", st["n_queue"], "of", nq, "with BF16 tiles", flush=True)
    assert st["n_queue"] < nq


@pytest.mark.parametrize("metric,precision", [(S.COSINE, S.BF16), (S.COSINE, S.FP32), (S.CHI2, S.BF16),
                                              (S.CHI2, S.FP32)])
def test_sampled_tile_error_within_band(ctx, metric, precision):
    """Candidate pairs sampled from the index: each pair is a real candidate,
    the FP64 value matches numpy's, and the tiles' arithmetic stays within the
    error band the eps decision relies on."""
    n, d = 6000, 784
    if metric == S.COSINE:
        X, _ = datagen.mixture(n, d, sigma=0.3, seed=211)
        eps, ker = 0.3, None
    else:
        X, _ = datagen.dirichlet(n, d, seed=211)
        eps, ker = 0.47, 1024
    p, bits, R = run_index(ctx, X, metric=metric, K=10, M=100, n_proj=1024, ker=ker, precision=precision)
    ctx.find_core(eps, 10)
    band = ctx.stats()["band"]
    a, b, fast, exact = ctx.sample_pair_dist(100000, seed=7)
    err = np.abs(fast.astype(np.float64) - exact)
    print(f"metric {metric} precision {precision}: max |fast - exact| {err.max():.3e}, band {band:.3e}",
          flush=True)
    assert err.max() <= band
    assert (a != b).all() and a.min() >= 0 and a.max() < n and b.min() >= 0 and b.max() < n
    tk, tm = ctx.get_index()
    K, M = 10, 100
    Xt = ctx.get_transformed().astype(np.float64)
    for j in range(0, 100000, 997):
        lists = [tm[tk[a[j], k], :M] for k in range(K)] + [tm[tk[a[j], K + k], M:] for k in range(K)]
        assert b[j] in np.concatenate(lists)
        x, y = Xt[a[j]], Xt[b[j]]
        if metric == S.COSINE:
            ref = 1.0 - x @ y
        elif precision == S.BF16:
            ref = 1.0 - np.sqrt(x) @ np.sqrt(y)
        else:
            ref = O.dist64(Xt[a[j]].astype(np.float32), Xt[b[j]].astype(np.float32), S.CHI2)
        assert abs(exact[j] - ref) < 1e-12


def test_device_dirichlet_generator(ctx):
    """Generator kind 1: nonnegative rows summing to one, reproducible from the
    seed, with the mass of each row concentrated on a few bins."""
    n, d = 4000, 784
    ctx.set_params(n, d, n_proj=256, topK=5, topM=20, metric=S.CHI2, ker_n_features=1024)
    ctx.generate(n, d, 1004, 0.1, kind=1)
    X = ctx.copy_input(0, n)
    ctx.generate(n, d, 1004, 0.1, kind=1)
    assert np.array_equal(X, ctx.copy_input(0, n))
    assert (X >= 0).all() and np.allclose(X.sum(1), 1.0, atol=1e-5)
    # 20 boosted bins (alpha 2) against 764 at alpha 0.05: expected mass 40/78.2
    top = np.sort(X, axis=1)[:, -20:].sum(1)
    print("mean top-20 mass", top.mean(), flush=True)
    assert top.mean() > 0.5
    ctx.generate(n, d, 1005, 0.1, kind=1)
    assert not np.array_equal(X, ctx.copy_input(0, n))


def test_sdbscan_eps_sweep_reuses_index(ctx):
    X, _ = datagen.mixture(6000, 128, k=6, sigma=0.3, seed=51)
    p, bits, R = run_index(ctx, X, K=8, M=60, n_proj=256)
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    for eps in (0.1, 0.2, 0.3, 0.4):
        core, deg = ctx.find_core(eps, 20)
        lab, nc = ctx.form_clusters()
        oc, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, eps, 20, mode=1)
        olab, onc = O.form_cluster(oc, ptr, idx)
        print("eps", eps, "clusters", nc, "cores", int(core.sum()), flush=True)
        assert np.array_equal(deg, od) and np.array_equal(core, oc)
        assert np.array_equal(lab, olab) and nc == onc


def test_optics_core_dist_and_reachability(ctx):
    X, _ = datagen.mixture(5000, 128, k=5, sigma=0.3, seed=61)
    p, bits, R = run_index(ctx, X, K=8, M=60, n_proj=256)
    eps, minPts = 0.3, 15
    core, deg = ctx.find_core(eps, minPts)
    ptr, idx, dist = ctx.get_neighbors()
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    oc, od, (optr, oidx, odd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, eps, minPts, mode=1)
    assert np.array_equal(ptr.astype(np.int32), optr)
    assert np.array_equal(idx, oidx)
    assert np.max(np.abs(dist - odd)) < 1e-5
    cd = ctx.core_dist(minPts)
    occ, ocd = O.core_dist(minPts, ptr.astype(np.int32), dist)
    assert np.array_equal(occ, core)
    assert np.array_equal(cd, ocd)
    order, reach = O.optics(core, cd, ptr.astype(np.int32), idx, dist)
    dreach = ctx.optics_reach(order)
    print("optics: unreached", int((reach == -2147483648.0).sum()), flush=True)
    assert np.array_equal(dreach, reach)
    # the ordering built by the library over the device neighbourhoods
    dorder, dreach2 = ctx.optics_order(minPts)
    assert np.array_equal(dorder, order)
    assert np.array_equal(dreach2, reach)


@pytest.mark.parametrize("metric", [S.COSINE, S.CHI2])
def test_optics_mirrored_neighbour_distances(ctx, metric):
    """Rows of more than 128 features: the neighbourhood distance pass computes
    each symmetric pair once and copies the other entry from its mirror.  The
    distances must equal the oracle's lane-order FP32 values bit for bit,
    every entry (a, c) must hold the bits of (c, a), and the walk over them
    must give the oracle's ordering and reachability."""
    if metric == S.COSINE:
        X = datagen.sphere_mixture(2500, 784, 12, seed=2171, spread=0.6)[0]
        p, bits, R = run_index(ctx, X, K=5, M=40, n_proj=256)
        eps = 0.35
    else:
        X, _ = datagen.dirichlet(2500, 300, k=5, seed=2173)
        p, bits, R = run_index(ctx, X, metric=S.CHI2, K=5, M=40, n_proj=256, ker=1024)
        eps = 0.3
    minPts = 10
    core, deg = ctx.find_core(eps, minPts)
    ptr, idx, dist = ctx.get_neighbors()
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    _, _, (optr, oidx, odd), _, _ = O.find_core(Xt, metric, tk, tm, eps, minPts, mode=3)
    n = X.shape[0]
    print(f"metric {metric} d={X.shape[1]} neighbourhoods: entries {int(ptr[-1])}, core {int(core.sum())}", flush=True)
    assert ptr[-1] > 10 * n
    assert np.array_equal(ptr.astype(np.int32), optr) and np.array_equal(idx, oidx)
    assert np.array_equal(dist.view(np.uint32), odd.view(np.uint32))
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
    key = rows * n + idx
    mirror = np.searchsorted(key, idx.astype(np.int64) * n + rows)
    assert np.array_equal(key[mirror], idx.astype(np.int64) * n + rows)
    assert np.array_equal(dist.view(np.uint32), dist[mirror].view(np.uint32))
    cd = ctx.core_dist(minPts)
    order, reach = O.optics(core, cd, ptr.astype(np.int32), idx, dist)
    dorder, dreach = ctx.optics_order(minPts)
    assert np.array_equal(dorder, order) and np.array_equal(dreach, reach)


def test_find_core_minpts_reuses_distances(ctx):
    """A second find_core at the same eps only moves the core test; its
    results equal a fresh index's."""
    X, _ = datagen.mixture(4000, 128, k=5, sigma=0.3, seed=141)
    print("minPts reuse n=4000 d=128", flush=True)
    run_index(ctx, X, K=6, M=40, n_proj=256)
    eps = 0.3
    c1, d1 = ctx.find_core(eps, 12)
    c2, d2 = ctx.find_core(eps, 25)
    lab2, nc2 = ctx.form_clusters()
    assert np.array_equal(d1, d2)
    assert np.array_equal(c2, (d1 >= 24).astype(np.uint8))
    assert 0 < int(c2.sum()) < int(c1.sum())
    run_index(ctx, X, K=6, M=40, n_proj=256)
    c3, d3 = ctx.find_core(eps, 25)
    lab3, nc3 = ctx.form_clusters()
    assert np.array_equal(c2, c3) and np.array_equal(d2, d3)
    assert np.array_equal(lab2, lab3) and nc2 == nc3


def test_tile_rows_autotune_keeps_results(ctx):
    X, _ = datagen.mixture(6000, 784, seed=151)
    print("tile rows autotune n=6000 d=784", flush=True)
    for precision in (S.FP32, S.BF16):
        run_index(ctx, X, K=5, M=50, precision=precision)
        core0, deg0 = ctx.find_core(0.25, 20)
        lab0, nc0 = ctx.form_clusters()
        best, ms = ctx.tune_tile_rows(0.25, [256, 1024, 4096])
        assert best in (256, 1024, 4096) and ms > 0
        core1, deg1 = ctx.find_core(0.25, 20)
        lab1, nc1 = ctx.form_clusters()
        assert np.array_equal(core0, core1) and np.array_equal(deg0, deg1)
        assert np.array_equal(lab0, lab1) and nc0 == nc1
        ctx.set_tile_rows(0)


def test_batch_matches_per_dataset_runs(ctx):
    """sdb_batch_run (3 streams, autotuned tiles) gives every dataset's labels,
    cluster counts, OPTICS ordering and reachability of one-by-one runs."""
    shapes = [(2500, 32), (4000, 128), (3000, 784), (6000, 32), (1500, 128)]
    data = [datagen.mixture(n, d, k=4, sigma=0.3, seed=170 + i)[0] for i, (n, d) in enumerate(shapes)]
    eps, minPts = [0.3, 0.2], [10, 20]
    print("batch 5 datasets x 2 eps x 2 minPts", flush=True)
    res, wall = S.batch_run(data, eps, minPts, optics=True, n_streams=3, n_proj=256, topK=5, topM=40,
                            hd3=O.hd3_bits)
    assert wall > 0
    for X, r in zip(data, res):
        assert r["status"] == 0 and r["tile_rows"] >= 0
        run_index(ctx, X, K=5, M=40, n_proj=256)
        for ie, e in enumerate(eps):
            for im, mp in enumerate(minPts):
                ctx.find_core(e, mp)
                lab, nc = ctx.form_clusters()
                assert np.array_equal(r["labels"][ie, im], lab) and r["n_clusters"][ie, im] == nc
        for im, mp in enumerate(minPts):
            ctx.find_core(max(eps), mp)
            o, rr = ctx.optics_order(mp)
            assert np.array_equal(r["ordering"][im], o) and np.array_equal(r["reach"][im], rr)


def test_label_noise_matches_oracle(ctx):
    X, _ = datagen.mixture(6000, 128, k=6, sigma=0.35, seed=91)
    print("label noise n=6000 d=128 K=8 M=50", flush=True)
    p, bits, R = run_index(ctx, X, K=8, M=50, n_proj=256)
    core, deg = ctx.find_core(0.33, 20)
    lab, nc = ctx.form_clusters()
    noise = int((lab == -1).sum())
    print("noise points", noise, "cores", int(core.sum()), flush=True)
    assert noise > 50 and core.sum() > 50
    cores = np.where(core)[0].astype(np.int32)
    sampled = cores[np.random.default_rng(0).random(len(cores)) < 0.3]
    glab = ctx.label_noise(sampled)
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    E = O.embed(Xt[sampled], S.COSINE, p.ker_n_features, p.ker_intervalSampling, R)
    cproj = O.project(E, bits, p.fhtDim, p.n_proj)
    olab = O.label_noise(cproj, tk, lab, sampled)
    assert np.array_equal(glab, olab)
    assert (glab == -1).sum() == 0


def test_dropin_module_cluster_noise(ctx):
    m = S.module()
    X, _ = datagen.mixture(4000, 128, k=5, sigma=0.35, seed=93)
    labs = {}
    for cn in (0, 1, 2):
        db = m.sDbscan(4000, 128)
        db.set_params(n_proj=256, topK=8, topM=50, distance="Cosine", clusterNoise=cn, samplingProb=0.05,
                      random_seed=42)
        db.fit_sDbscan(X.T, 0.33, 20)
        labs[cn] = np.array(db.labels_)
    print("noise with clusterNoise 0/1/2:", [int((labs[c] == -1).sum()) for c in (0, 1, 2)], flush=True)
    assert (labs[0] == -1).sum() > 0 and (labs[1] == -1).sum() == 0
    assert np.array_equal(labs[0], labs[2])           # the heuristic changes nothing (no kept neighbourhoods)
    keep = labs[0] >= 0
    assert np.array_equal(labs[1][keep], labs[0][keep])


def test_dropin_module_text_loader(ctx, tmp_path):
    m = S.module()
    X, _ = datagen.mixture(2000, 64, k=4, sigma=0.3, seed=95)
    path = tmp_path / "points.txt"
    np.savetxt(path, X)                     # one point per line, exact float32 round trip
    runs = []
    for use_file in (False, True):
        db = m.sDbscan(2000, 64)
        db.set_params(n_proj=256, topK=5, topM=30, distance="Cosine", random_seed=7)
        if use_file:
            db.load_fit_sDbscan(str(path), 0.3, 10)
        else:
            db.fit_sDbscan(X.T, 0.3, 10)
        runs.append((np.array(db.labels_), db.n_clusters_))
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]


def std_uniform_ints(seed, N, count):
    """std::default_random_engine (minstd_rand0) + libstdc++
    uniform_int_distribution<int>(0, N - 1), published algorithms."""
    m = 2147483647
    x = seed % m or 1
    urange = (m - 1) - 1
    scaling = urange // N
    past = N * scaling
    out = []
    while len(out) < count:
        while True:
            x = (16807 * x) % m
            r = x - 1
            if r < past:
                break
        out.append(r // scaling)
    return np.array(out, np.int32)


def sng_windows(seq, n, S):
    """Window end per point: the shortest prefix holding S draws other than
    the point (sng_findCorePoints' while loop)."""
    win = np.full(n, S, np.int32)
    for v in np.unique(seq):
        pos = np.where(seq == v)[0]
        c = 0
        while c < len(pos) and pos[c] < S + c:
            c += 1
        win[v] = S + c
    return win


def test_sng_exact_matches_oracle(ctx):
    n = 3000
    X, _ = datagen.mixture(n, 64, k=5, sigma=0.3, seed=101)
    print("sng exact n=3000 d=64", flush=True)
    ctx.set_params(n, 64, n_proj=64, topK=4, topM=20, metric=S.COSINE)
    ctx.load(X)
    core, deg = ctx.sng_find_core(np.arange(n), 0.2, 15, both_ways=False)
    lab, nc = ctx.form_clusters()
    Xt = ctx.get_transformed()
    oc, od, (ptr, idx, dd) = O.sng_find_core(Xt, S.COSINE, np.arange(n), 0.2, 15, both_ways=False)
    olab, onc = O.form_cluster(oc, ptr, idx)
    print("cores", int(core.sum()), "clusters", nc, flush=True)
    assert np.array_equal(deg, od) and np.array_equal(core, oc)
    assert np.array_equal(lab, olab) and nc == onc


@pytest.mark.parametrize("precision", [S.FP32, S.BF16])
@pytest.mark.parametrize("metric", [S.COSINE, S.CHI2])
def test_sng_sampled_matches_oracle(ctx, precision, metric):
    n, S_ = 5000, 500
    if metric == S.COSINE:
        X, _ = datagen.mixture(n, 128, k=5, sigma=0.3, seed=103)
        d, eps = 128, 0.3
    else:
        X, _ = datagen.dirichlet(n, 100, k=5, seed=103)
        d, eps = 100, 0.25
    seq = std_uniform_ints(11, n, S_ + 64)
    seq[10] = seq[20] = 7            # repeated draws
    seq[30] = 30                     # a point drawing itself
    win = sng_windows(seq, n, S_)
    print(f"sng sampled n=5000 d={d} samples=500 metric {metric} precision {precision}", flush=True)
    ctx.set_params(n, d, n_proj=128, topK=4, topM=20, metric=metric, ker_n_features=max(d, 128),
                   precision=precision)
    ctx.load(X)
    core, deg = ctx.sng_find_core(seq, eps, 12, win_end=win)
    lab, nc = ctx.form_clusters()
    Xt = ctx.get_transformed()
    oc, od, (ptr, idx, dd) = O.sng_find_core(Xt, metric, seq, eps, 12, win_end=win)
    olab, onc = O.form_cluster(oc, ptr, idx)
    print("cores", int(core.sum()), "clusters", nc, "queue", ctx.stats()["n_queue"], flush=True)
    assert np.array_equal(deg, od) and np.array_equal(core, oc)
    assert np.array_equal(lab, olab) and nc == onc


def test_sng_optics_matches_oracle(ctx):
    n = 4000
    X, _ = datagen.mixture(n, 96, k=4, sigma=0.3, seed=105)
    seq = std_uniform_ints(5, n, 400)
    ctx.set_params(n, 96, n_proj=128, topK=4, topM=20, metric=S.COSINE)
    ctx.load(X)
    eps, minPts = 0.3, 10
    core, deg = ctx.sng_find_core(seq, eps, minPts, dedup=True)
    ptr, idx, dist = ctx.get_neighbors()
    Xt = ctx.get_transformed()
    oc, od, (optr, oidx, odd) = O.sng_find_core(Xt, S.COSINE, seq, eps, minPts, dedup=True)
    assert np.array_equal(core, oc) and np.array_equal(deg, od)
    assert np.array_equal(ptr.astype(np.int32), optr) and np.array_equal(idx, oidx)
    assert np.max(np.abs(dist - odd)) < 1e-5
    cd = ctx.core_dist(minPts)
    occ, ocd = O.core_dist(minPts, ptr.astype(np.int32), dist)
    assert np.array_equal(cd, ocd)
    order, reach = O.optics(core, cd, ptr.astype(np.int32), idx, dist)
    assert np.array_equal(ctx.optics_reach(order), reach)


def test_dropin_module_sng(ctx):
    m = S.module()
    n = 3000
    X, _ = datagen.mixture(n, 64, k=4, sigma=0.3, seed=107)
    db = m.sDbscan(n, 64)
    db.set_params(distance="Cosine", samplingProb=0.1, random_seed=3)
    db.fit_sngDbscan(X.T, 0.25, 10)
    # the same draw sequence through the C ABI
    S_ = int(np.ceil(1.0 * n * float(np.float32(0.1))))   # ceil(1.0 * n * prob) with prob a float
    seq = std_uniform_ints(3, n, S_ + 64)
    win = sng_windows(seq, n, S_)
    ctx.set_params(n, 64, n_proj=64, topK=4, topM=20, metric=S.COSINE)
    ctx.load(X)
    ctx.sng_find_core(seq, 0.25, 10, win_end=win)
    lab, nc = ctx.form_clusters()
    assert np.array_equal(np.array(db.labels_), lab) and db.n_clusters_ == nc
    db.fit_sngOptics(X.T, 0.25, 10)
    assert sorted(db.ordering_) == list(range(n))


def test_dropin_module_fit(ctx):
    m = S.module()
    X, _ = datagen.mixture(4000, 784, seed=71)
    db = m.sDbscan(4000, 784)
    db.set_params(n_proj=1024, topK=5, topM=50, distance="Cosine", random_seed=42)
    db.fit_sDbscan(X.T, 0.25, 50)
    # same seed through the C ABI
    run_index(ctx, X, K=5, M=50, seed=42)
    ctx.find_core(0.25, 50)
    lab, nc = ctx.form_clusters()
    assert np.array_equal(np.array(db.labels_), lab) and db.n_clusters_ == nc
    db.fit_sOptics(X.T, 0.25, 50)
    assert sorted(db.ordering_) == list(range(4000))
    assert len(db.reachability_) == 4000


def test_error_paths(ctx):
    X = np.abs(np.random.default_rng(0).standard_normal((500, 20))).astype(np.float32)
    Xn = X.copy()
    Xn[10, 3] = -1.0
    with pytest.raises(S.SdbError) as e:
        run_index(ctx, Xn, metric=S.CHI2, K=4, M=20, n_proj=64, ker=64)
    assert e.value.code == 4
    # the same checks where the transform is fused into the index passes
    with pytest.raises(S.SdbError) as e:
        run_index(ctx, Xn, metric=S.CHI2, K=4, M=20, n_proj=64, ker=32)
    assert e.value.code == 4
    Xz = X.copy()
    Xz[5] = 0.0
    with pytest.raises(S.SdbError) as e:
        run_index(ctx, Xz, metric=S.COSINE, K=4, M=20, n_proj=64)
    assert e.value.code == 5
    with pytest.raises(S.SdbError) as e:
        run_index(ctx, Xz, metric=S.JS, K=4, M=20, n_proj=64, ker=32)
    assert e.value.code == 5
    with pytest.raises(S.SdbError) as e:
        ctx.set_params(500, 20, n_proj=64, topK=100, topM=20)
    assert e.value.code == 1
    # parameters of other dimensions drop the loaded input (SDB_ESTATE until
    # the next load); the same dimensions keep it
    run_index(ctx, X, metric=S.COSINE, K=4, M=20, n_proj=64)
    p = ctx.set_params(400, 20, n_proj=64, topK=4, topM=20)
    bits = O.hd3_bits(3 * p.fhtDim, 42)
    with pytest.raises(S.SdbError) as e:
        ctx.build_index(bits)
    assert e.value.code == 6
    ctx.load(X[:400])
    ctx.build_index(bits)
    ctx.set_params(400, 20, n_proj=64, topK=4, topM=20)
    ctx.build_index(bits)
    tk, _ = ctx.get_index()
    assert tk.shape[0] == 400
