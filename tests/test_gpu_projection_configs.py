"""Device HD3 projections read back through sdb_get_projections, and the
L2 (random Fourier features) and many-dataset batch workloads at their own
parameters, against the oracle.

Projections: bit-equal (as IEEE values; +0 == -0) to the oracle's, which is
bit-pinned to the reference's compiled fht_float chain, wherever the
embedding is exact arithmetic (Cosine, Chi2 square-root features); within
1e-5 relative (per point, against its largest projection) where the
embedding goes through cos / sin / log (RFF, sampled Chi2 / JS features),
whose device and host libm values may differ in the last ulp.
"""
import os

import numpy as np
import pytest

import sdbscan_amd as S
from oracle import oracle as O
from sdbscan_amd import datagen

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def setup(ctx, X, metric=S.COSINE, n_proj=1024, K=5, M=50, ker=None, sigma=1.0, L=0.4, seed=42,
          precision=S.FP32, bits=None):
    n, d = X.shape
    p = ctx.set_params(n, d, n_proj=n_proj, topK=K, topM=M, metric=metric, ker_n_features=ker,
                       ker_sigma=sigma, ker_intervalSampling=L, precision=precision)
    if bits is None:
        bits = O.hd3_bits(3 * p.fhtDim, seed)
    R = None
    if metric in (S.L1, S.L2):
        half = p.ker_n_features // 2
        R = (O.gauss if metric == S.L2 else O.cauchy)(half, d, 0.0, 1.0 / sigma, seed)
    ctx.load(X)
    ctx.build_index(bits, R)
    return p, bits, R


def oracle_proj(ctx, p, bits, R, metric):
    Xt = ctx.get_transformed()
    E = O.embed(Xt, metric, p.ker_n_features, p.ker_intervalSampling, R)
    return O.project(E, bits, p.fhtDim, p.n_proj), Xt   # P x n


def test_projections_reproduce_the_golden_hd3_chains(ctx):
    """The golden HD3 chains (3 x {sign flip, fht_float}, written by the
    reference's own FFHT) fed through build_index: rows whose Cosine
    transform leaves them unchanged give the golden outputs exactly."""
    g = np.load(os.path.join(GOLD, "fwht_golden.npz"))
    hin, hout, bits = g["hd3_in"], g["hd3_out"], g["hd3_bits_seed42"]
    filler, _ = datagen.mixture(2000, 784, seed=3)
    X = np.concatenate([hin, filler]).astype(np.float32)
    print("golden hd3 chains through build_index, n=2006", flush=True)
    p, _, _ = setup(ctx, X, bits=bits)
    Xt = ctx.get_transformed()
    same = [i for i in range(len(hin)) if np.array_equal(Xt[i].view(np.uint32), hin[i].view(np.uint32))]
    assert len(same) >= 3, same
    dev = ctx.get_projections(np.arange(len(hin)))
    for i in same:
        assert np.array_equal(dev[i], hout[i]), i          # as values
    # every row, against the oracle's chain over the device's transformed rows
    proj, _ = oracle_proj(ctx, p, bits, None, S.COSINE)
    assert np.array_equal(dev, proj[:, :len(hin)].T)


def test_projections_bit_equal_at_c0(ctx):
    """4,000 points of the 20,000 x 784 cosine workload: every projection
    equals the oracle's (ordered listing and a shuffled one)."""
    X, _ = datagen.mixture(4000, 784, seed=1000)
    print("projections cosine n=4000 d=784 P=1024", flush=True)
    p, bits, R = setup(ctx, X, K=5, M=50)
    proj, _ = oracle_proj(ctx, p, bits, R, S.COSINE)
    dev = ctx.get_projections(np.arange(4000))
    assert np.array_equal(dev, proj.T)
    ids = np.random.default_rng(1).permutation(4000)[:777]
    assert np.array_equal(ctx.get_projections(ids), proj.T[ids])
    # the index is the selection over these projections
    tk, tm = ctx.get_index()
    assert all(set(a) == set(b) for a, b in zip(tk, O.topk(proj, p.topK)))


def rel_close(dev, ref, tol=1e-5):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    return np.abs(dev - ref) <= tol * scale


@pytest.mark.parametrize("metric,d,ker,exact", [
    (S.CHI2, 784, 1024, True),    # square-root features, fused transform
    (S.CHI2, 100, 700, False),    # 3 sampled components
    (S.JS, 100, 700, False),
    (S.JS, 300, 512, False),      # square-root features with the JS constant
    (S.L2, 51, 1024, False),      # random Fourier features
    (S.L1, 40, 512, False),
])
def test_projections_other_embeddings(ctx, metric, d, ker, exact):
    if metric in (S.L1, S.L2):
        X, _ = datagen.blobs(3000, d, k=5, seed=77)
        kw = dict(sigma=6.0)
    else:
        X, _ = datagen.dirichlet(3000, d, k=5, seed=77)
        kw = {}
    print(f"projections metric {metric} d={d} ker={ker}", flush=True)
    p, bits, R = setup(ctx, X, metric=metric, ker=ker, n_proj=1024, K=6, M=40, **kw)
    proj, _ = oracle_proj(ctx, p, bits, R, metric)
    dev = ctx.get_projections(np.arange(3000))
    ok = rel_close(dev, proj.T)
    print("max rel", float((np.abs(dev - proj.T) / np.abs(proj.T).max(1, keepdims=True)).max()), flush=True)
    assert ok.all()
    if exact:
        assert np.array_equal(dev, proj.T)
    # staged: the device index is the exact selection over the device's own projections
    tk, tm = ctx.get_index()
    assert all(set(a) == set(b) for a, b in zip(tk[:, :6], O.topk(dev.T, 6)[:, :6]))
    assert all(set(a) == set(b) for a, b in zip(tm, O.topm(dev.T, 40)))


def test_l2_rff_workload_parameters(ctx):
    """The 1.8M x 51 L2 workload's parameters on 20,000 of its points: 12
    anisotropic blobs + 5% noise, RFF to 1024 dims with sigma = median
    pairwise L2 of a 10k sample, eps = the L2 radius of embedded cosine
    distance 0.3, topK 10, topM 100 (two candidate columns per lane in the
    51-dim VALU tiles), minPts 50.  Projections within 1e-5 of the oracle's;
    index = the selection over them; degrees, cores and labels equal the
    oracle's on the device index."""
    X, _ = datagen.blobs(20000, 51, k=12, noise=0.05, seed=1001)
    sigma = datagen.median_pairwise_l2(X, 10000, seed=1001)
    eps = datagen.l2_eps_for_embedded_cosine(sigma, 0.3)
    print(f"L2 RFF n=20000 d=51 sigma={sigma:.3f} eps={eps:.3f} K=10 M=100 minPts=50", flush=True)
    p, bits, R = setup(ctx, X, metric=S.L2, ker=1024, sigma=sigma, K=10, M=100)
    proj, Xt = oracle_proj(ctx, p, bits, R, S.L2)
    dev = ctx.get_projections(np.arange(20000))
    assert rel_close(dev, proj.T).all()
    tk, tm = ctx.get_index()
    assert all(set(a) == set(b) for a, b in zip(tk[:, :10], O.topk(dev.T, 10)[:, :10]))
    assert all(set(a) == set(b) for a, b in zip(tk[:, 10:], O.topk(dev.T, 10)[:, 10:]))
    assert all(set(a) == set(b) for a, b in zip(tm, O.topm(dev.T, 100)))
    core, deg = ctx.find_core(eps, 50)
    lab, nc = ctx.form_clusters()
    st = ctx.stats()
    oc, od, (ptr, idx, dd), near, _ = O.find_core(Xt, S.L2, tk, tm, eps, 50, mode=1)
    olab, onc = O.form_cluster(oc, ptr, idx)
    print("clusters", nc, "cores", int(core.sum()), "queue", st["n_queue"], "near", near, flush=True)
    assert np.array_equal(deg, od) and np.array_equal(core, oc)
    assert np.array_equal(lab, olab) and nc == onc
    assert 0 < int(core.sum()) < 20000


def test_batch_sweep_matches_oracle(ctx):
    """sdb_batch_run over six unit-sphere mixtures (d in {32, 128, 784}, n
    2,000-6,000, 3-30 components, spreads chosen so that the sweep crosses
    from all-noise to merged clusters), 8 eps values in [0.1, 0.4] x minPts
    {10, 50}, sOPTICS at the largest eps, topK 10, topM 50, FP32: every label
    set and cluster count, and every OPTICS ordering and reachability, equal
    the oracle's (index, FP64 decisions, formCluster, and formOptics_scikit
    over the neighbourhood distances in the device's order)."""
    rng = np.random.default_rng(1003)
    shapes = [(6000, 32, 25, 0.5), (5000, 128, 7, 0.6), (2500, 784, 12, 0.6), (4000, 32, 3, 0.7),
              (3000, 128, 30, 0.55), (2000, 784, 5, 0.65)]
    data = [datagen.sphere_mixture(n, d, k, seed=2100 + i, spread=sp)[0] for i, (n, d, k, sp) in enumerate(shapes)]
    eps = np.sort(rng.uniform(0.1, 0.4, 8)).astype(np.float32)
    minPts = [10, 50]
    print("batch 6 sphere mixtures x 8 eps x 2 minPts + optics, vs oracle", flush=True)
    res, wall = S.batch_run(data, eps, minPts, optics=True, n_streams=3, n_proj=1024, topK=10, topM=50,
                            hd3=O.hd3_bits)
    seen = set()
    for (n, d, k, sp), X, r in zip(shapes, data, res):
        assert r["status"] == 0, r["error"]
        Xt, rc = O.transform(X, S.COSINE)
        assert rc == 0
        D = S.fht_dim(1024, d)
        bits = O.hd3_bits(3 * D, 42)
        proj = O.project(O.embed(Xt, S.COSINE, d, 0.4, None), bits, D, 1024)
        tk, tm = O.topk(proj, 10), O.topm(proj, 50)
        for ie, e in enumerate(eps):
            # the neighbourhoods of eps do not depend on minPts
            _, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, float(e), 1, mode=3)
            for im, mp in enumerate(minPts):
                oc = (od >= mp - 1).astype(np.uint8)
                olab, onc = O.form_cluster(oc, ptr, idx)
                assert np.array_equal(r["labels"][ie, im], olab), (n, d, e, mp)
                assert r["n_clusters"][ie, im] == onc
                seen.add((int(onc), int((olab == -1).sum()) > 0))
        for im, mp in enumerate(minPts):
            # ptr / idx / dd of the largest eps (the last of the sweep)
            oc = (od >= mp - 1).astype(np.uint8)
            occ, ocd = O.core_dist(max(mp, 2), ptr, dd)
            assert np.array_equal(occ, oc)
            order, reach = O.optics(oc, ocd, ptr, idx, dd)
            assert np.array_equal(r["ordering"][im], order), (n, d, mp)
            assert np.array_equal(r["reach"][im], reach), (n, d, mp)
        print(f"dataset n={n} d={d} k={k} ok, {r['ms']:.0f} ms, tile rows {r['tile_rows']}", flush=True)
    assert len(seen) > 8   # the sweep moves through different clusterings


def test_entry_count_guard(ctx):
    """set_params refuses n * 2 * topK >= 2^31 (the entries are counted and
    grouped in signed 32-bit offsets) before any allocation, and accepts one
    point fewer."""
    n = (1 << 31) // (2 * 16)
    with pytest.raises(S.SdbError) as e:
        ctx.set_params(n, 8, n_proj=64, topK=16, topM=10)
    assert e.value.code == 7
    ctx.set_params(n - 1, 8, n_proj=64, topK=16, topM=10)


def test_set_params_drops_transformed_rows_of_another_metric(ctx):
    """Cosine index, then set_params to Chi2 on the same input: the sample
    neighbourhoods are computed on the Chi2-transformed rows (equal to the
    oracle's), not on the Cosine rows left from the index."""
    X, _ = datagen.dirichlet(3000, 100, k=4, seed=12)
    setup(ctx, X, metric=S.COSINE, n_proj=256, K=5, M=40)
    ctx.set_params(3000, 100, n_proj=256, topK=5, topM=40, metric=S.CHI2, ker_n_features=256)
    seq = np.random.default_rng(2).integers(0, 3000, 600).astype(np.int32)
    core, deg = ctx.sng_find_core(seq, 0.3, 5)
    Xo, rc = O.transform(X, S.CHI2)
    assert np.array_equal(ctx.get_transformed().view(np.uint32), Xo.view(np.uint32))
    oc, od, _ = O.sng_find_core(Xo, S.CHI2, seq, 0.3, 5)
    assert np.array_equal(deg, od) and np.array_equal(core, oc)


def one_hot_rows(n, d, n_hot, seed):
    X, _ = datagen.mixture(n, d, k=3, seed=seed)
    hot = np.random.default_rng(seed).integers(0, d, n_hot)
    X[:n_hot] = 0.0
    X[np.arange(n_hot), hot] = 1.0   # +-1 projections: every top-K boundary is tied
    return X


def test_tied_topk_replayed_in_batches(ctx, monkeypatch):
    """More tied points than the values kept per pass (cap lowered to 64 for
    the test): the rest are projected again and replayed batch by batch;
    the index equals the oracle's heaps."""
    monkeypatch.setenv("SDB_TIE_CAP", "64")
    X = one_hot_rows(3000, 96, 700, 41)
    print("tied top-K in batches of 64: n=3000, 700 one-hot rows", flush=True)
    p, bits, R = setup(ctx, X, n_proj=128, K=8, M=30)
    st = ctx.stats()
    assert st["n_tie_points"] > 3 * 64
    tk, tm = ctx.get_index()
    proj, Xt = oracle_proj(ctx, p, bits, R, S.COSINE)
    tko = O.topk(proj, 8)
    assert np.array_equal(np.sort(tk, axis=1), np.sort(tko, axis=1))


def test_more_than_65536_tied_points_match_the_oracle(ctx):
    """110,000 one-hot points among 112,000, most with a tied top-K boundary
    (more than the 65,536 whose values one pass keeps): the index (the
    reference's heap choices among equal values), degrees, cores and labels
    equal the oracle's."""
    X = one_hot_rows(112000, 64, 110000, 43)
    print("tied top-K: n=112000 d=64, 110000 one-hot rows, P=128 K=4 M=20", flush=True)
    p, bits, R = setup(ctx, X, n_proj=128, K=4, M=20)
    st = ctx.stats()
    print("tie points", st["n_tie_points"], "tie rows", st["n_tie_rows"], flush=True)
    assert st["n_tie_points"] > 65536
    tk, tm = ctx.get_index()
    proj, Xt = oracle_proj(ctx, p, bits, R, S.COSINE)
    tko, tmo = O.topk(proj, 4), O.topm(proj, 20)
    # heap choices among equal values: the same sets (the order inside a
    # list is the heap's internal order)
    assert np.array_equal(np.sort(tk[:, :4], axis=1), np.sort(tko[:, :4], axis=1))
    assert np.array_equal(np.sort(tk[:, 4:], axis=1), np.sort(tko[:, 4:], axis=1))
    assert np.array_equal(np.sort(tm[:, :20], axis=1), np.sort(tmo[:, :20], axis=1))
    assert np.array_equal(np.sort(tm[:, 20:], axis=1), np.sort(tmo[:, 20:], axis=1))
    core, deg = ctx.find_core(0.3, 10)
    lab, nc = ctx.form_clusters()
    oc, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, 0.3, 10, mode=1)
    olab, onc = O.form_cluster(oc, ptr, idx)
    assert np.array_equal(deg, od) and np.array_equal(core, oc)
    assert np.array_equal(lab, olab) and nc == onc


def oracle_cosine_index(X, n_proj, K, M, seed=42):
    Xt, rc = O.transform(X, S.COSINE)
    assert rc == 0
    D = S.fht_dim(n_proj, X.shape[1])
    bits = O.hd3_bits(3 * D, seed)
    proj = O.project(O.embed(Xt, S.COSINE, X.shape[1], 0.4, None), bits, D, n_proj)
    return Xt, O.topk(proj, K), O.topm(proj, M)


def test_dropin_module_sweep_and_optics_files_match_the_oracle(tmp_path):
    """The pybind module's test_sDbscan (an eps sweep on one index,
    src/sDbscan.cpp:980-1039) and fit_sOptics (:1711-1769) with an output
    prefix: every label file (one label per line) and the OPTICS file
    ("order reach" per line) carry the oracle's results, under the
    reference's file names."""
    m = S.module()
    X, _ = datagen.mixture(3000, 784, k=6, sigma=0.3, seed=77)
    out = str(tmp_path / "run")
    db = m.sDbscan(3000, 784)
    db.set_params(n_proj=1024, topK=5, topM=50, distance="Cosine", random_seed=42, output=out)
    eps0, rng_eps, n_tests, minPts = 0.25, 0.05, 3, 20
    print("module test_sDbscan + fit_sOptics, cosine 3000 x 784", flush=True)
    db.test_sDbscan(X.T, eps0, rng_eps, n_tests, minPts)
    Xt, tk, tm = oracle_cosine_index(X, 1024, 5, 50)
    for i in range(n_tests):
        e = float(np.float32(eps0 + 1.0 * i * np.float32(rng_eps)))
        name = f"{out}_Cosine_Eps_{int(np.round(np.float32(1000) * np.float32(e)))}_MinPts_{minPts}" \
               f"_KerFeatures_784_NumProj_1024_TopM_50_TopK_5"
        lab = np.loadtxt(name, dtype=np.int64)
        oc, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, e, minPts, mode=1)
        olab, onc = O.form_cluster(oc, ptr, idx)
        assert np.array_equal(lab, olab), e
    assert np.array_equal(np.array(db.labels_), olab)   # the last sweep step
    eps = 0.3
    db.fit_sOptics(X.T, eps, minPts)
    oc, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, eps, minPts, mode=3)
    occ, ocd = O.core_dist(minPts, ptr, dd)
    order, reach = O.optics(oc, ocd, ptr, idx, dd)
    assert np.array_equal(np.array(db.ordering_), order)
    assert np.array_equal(np.array(db.reachability_, np.float32), reach)
    name = f"{out}_Cosine_Eps_300_MinPts_{minPts}_NumEmbed_784_NumProj_1024_TopM_50_TopK_5"
    rows = [l.split() for l in open(name).read().splitlines()]
    assert len(rows) == 3000
    assert [int(r[0]) for r in rows] == order.tolist()
    # reach printed with the stream's default 6 significant digits
    assert all(abs(float(r[1]) - float(v)) <= 1e-5 * max(1.0, abs(float(v))) for r, v in zip(rows, reach))
