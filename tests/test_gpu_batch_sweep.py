"""The one-launch batched sweep (sdb_batch_sweep) against the oracle and
against the stream-pool batch runner (sdb_batch_run): labels, cluster counts
and sOPTICS outputs of every (dataset, eps, minPts) equal; the captured
cluster graphs give the same outputs on every replay."""
import numpy as np
import pytest

import sdbscan_amd as S
from oracle import oracle as O
from sdbscan_amd import datagen

pytestmark = pytest.mark.gpu

SHAPES = [(6000, 32, 25, 0.5), (5000, 128, 7, 0.6), (2500, 784, 12, 0.6), (4000, 32, 3, 0.7),
          (3000, 128, 30, 0.55), (2000, 784, 5, 0.65)]


def datasets():
    return [datagen.sphere_mixture(n, d, k, seed=2100 + i, spread=sp)[0] for i, (n, d, k, sp) in enumerate(SHAPES)]


def sweep_eps():
    return np.sort(np.random.default_rng(1003).uniform(0.1, 0.4, 8)).astype(np.float32)


def test_sweep_matches_oracle_and_batch_runner(monkeypatch):
    monkeypatch.setenv("SDB_GRAPH_REPLAYS", "3")   # every graph replayed three times, outputs compared
    data = datasets()
    eps = sweep_eps()[::-1].copy()   # unsorted order: outputs keep the caller's eps order
    minPts = [10, 50]
    print("one-launch sweep: 6 datasets x 8 eps x 2 minPts + optics", flush=True)
    res, wall, stages = S.batch_sweep(data, eps, minPts, optics=True, n_streams=3, n_proj=1024, topK=10, topM=50,
                                      hd3=O.hd3_bits)
    print("stages ms", {k: round(v, 1) for k, v in stages.items()}, "wall", round(wall, 1), flush=True)
    ref, _ = S.batch_run(data, eps, minPts, optics=True, n_streams=3, autotune=False, n_proj=1024, topK=10,
                         topM=50, hd3=O.hd3_bits)
    for (n, d, k, sp), X, r, rr in zip(SHAPES, data, res, ref):
        assert r["status"] == 0, r["error"]
        assert np.array_equal(r["labels"], rr["labels"]), (n, d)
        assert np.array_equal(r["n_clusters"], rr["n_clusters"]), (n, d)
        assert np.array_equal(r["ordering"], rr["ordering"]) and np.array_equal(r["reach"], rr["reach"])
    # the oracle, for two of the datasets (the sweep path itself, not only
    # agreement with the other runner)
    for i in (2, 4):
        X = data[i]
        Xt, rc = O.transform(X, S.COSINE)
        D = S.fht_dim(1024, X.shape[1])
        proj = O.project(O.embed(Xt, S.COSINE, X.shape[1], 0.4, None), O.hd3_bits(3 * D, 42), D, 1024)
        tk, tm = O.topk(proj, 10), O.topm(proj, 50)
        for ie, e in enumerate(eps):
            _, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, float(e), 1, mode=3)
            for im, mp in enumerate(minPts):
                olab, onc = O.form_cluster((od >= mp - 1).astype(np.uint8), ptr, idx)
                assert np.array_equal(res[i]["labels"][ie, im], olab) and res[i]["n_clusters"][ie, im] == onc
        _, od, (ptr, idx, dd), _, _ = O.find_core(Xt, S.COSINE, tk, tm, float(eps.max()), 1, mode=3)
        for im, mp in enumerate(minPts):
            oc = (od >= mp - 1).astype(np.uint8)
            _, ocd = O.core_dist(mp, ptr, dd)
            order, reach = O.optics(oc, ocd, ptr, idx, dd)
            assert np.array_equal(res[i]["ordering"][im], order) and np.array_equal(res[i]["reach"][im], reach)


def test_sweep_autotune_and_refusals():
    data = datasets()[:3]
    eps = sweep_eps()[:4]
    res, wall, stages = S.batch_sweep(data, eps, [10], optics=False, n_streams=2, autotune=True, n_proj=256,
                                      topK=5, topM=40, hd3=O.hd3_bits)
    ref, _ = S.batch_run(data, eps, [10], optics=False, n_streams=2, autotune=False, n_proj=256, topK=5, topM=40,
                         hd3=O.hd3_bits)
    for r, rr in zip(res, ref):
        assert r["tile_rows"] in (256, 1024, 2048, 4096)
        assert np.array_equal(r["labels"], rr["labels"])
    assert stages["autotune"] > 0
    with pytest.raises(S.SdbError) as e:
        S.batch_sweep(data[:1], np.linspace(0.1, 0.4, 17), [10], optics=False, n_proj=256, topK=5, topM=40)
    assert e.value.code == 7
