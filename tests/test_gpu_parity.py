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


def test_transform_bit_exact(ctx):
    X, _ = datagen.mixture(3000, 784, seed=7)
    X = X * np.float32(3.7)
    print("transform cosine n=3000 d=784", flush=True)
    run_index(ctx, X)
    Xg = ctx.get_transformed()
    Xo, rc = O.transform(X, S.COSINE)
    assert rc == 0
    assert np.array_equal(Xg.view(np.uint32), Xo.view(np.uint32))


def test_index_matches_oracle(ctx):
    X, _ = datagen.mixture(4000, 784, seed=11)
    print("index cosine n=4000 K=5 M=50", flush=True)
    p, bits, R = run_index(ctx, X, K=5, M=50)
    tk, tm = ctx.get_index()
    Xt = ctx.get_transformed()
    proj, tko, tmo = oracle_index(Xt, S.COSINE, p, bits, R)
    st = ctx.stats()
    print("stats", st, flush=True)
    assert np.array_equal(tk, tko)
    assert sets_equal(tm[:, :p.topM], tmo[:, :p.topM])
    assert sets_equal(tm[:, p.topM:], tmo[:, p.topM:])


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
