"""The command-line front end `sdbscan` (src/main.cpp:15-85) end to end on the
device: a text dataset read like loadtxtData (src/Utilities.cpp:60-95, which
transforms the rows once) and main's ten eps steps of fit_sDbscan and
fit_sngDbscan (each transforms them again), with --verbose and --output so
that every fit writes its label file under the reference's file names.  Every
label file must hold the oracle's labels for the same twice-transformed rows,
seed and eps."""
import math
import os
import subprocess

import numpy as np
import pytest

import sdbscan_amd as S
from oracle import oracle as O
from sdbscan_amd import datagen
from test_gpu_parity import sng_windows, std_uniform_ints

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sdbscan_amd", "sdbscan")


def lround(x):
    """std::lround of a float: half away from zero."""
    x = float(x)
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


@pytest.mark.parametrize("dist", ["Cosine", "Chi2"])
def test_cli_label_files_match_the_oracle(tmp_path, dist):
    n, K, M, P, minPts, seed = 2000, 5, 40, 256, 10, 7
    if dist == "Cosine":
        d, eps0, metric = 64, np.float32(0.2), S.COSINE
        X, _ = datagen.mixture(n, d, k=5, sigma=0.3, seed=211)
    else:
        # Chi2: kernel features default to n_features (square-root features),
        # interval sampling 0.4
        d, eps0, metric = 32, np.float32(0.3), S.CHI2
        X, _ = datagen.dirichlet(n, d, k=5, n_boost=6, seed=212)
    prob = np.float32(0.1)
    data = tmp_path / "data.txt"
    np.savetxt(data, X, fmt="%.9g")
    out = str(tmp_path / "run")
    cmd = [CLI, "--n_points", n, "--n_features", d, "--eps", f"{float(eps0):.9g}", "--minPts", minPts,
           "--topK", K, "--topM", M, "--n_proj", P, "--distance", dist, "--samplingProb", "0.1",
           "--random_seed", seed, "--verbose", "--output", out, "--dataset", data]
    print(f"sdbscan CLI, {dist} {n} x {d}, 10 eps steps", flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("sDBSCAN Wall Clock = ") == 10
    assert r.stdout.count("sngDBSCAN Wall Clock = ") == 10

    # the rows the fits see: the loader's transform, then the fit's own
    X1, rc = O.transform(X, metric)
    assert rc == 0
    Xt, rc = O.transform(X1, metric)
    assert rc == 0
    D = S.fht_dim(P, d)
    bits = O.hd3_bits(3 * D, seed)
    proj = O.project(O.embed(Xt, metric, d, 0.4, None), bits, D, P)
    tk, tm = O.topk(proj, K), O.topm(proj, M)

    S_ = int(math.ceil(1.0 * n * float(prob)))
    seq = std_uniform_ints(seed, n, S_ + 64)
    win = sng_windows(seq, n, S_)
    n_clustered = 0
    for i in range(10):
        e = np.float32(eps0 + np.float32(i) * np.float32(0.01))   # float arithmetic of main
        tag = f"{out}_{dist}_Eps_{lround(np.float32(1000) * e)}_MinPts_{minPts}"
        lab = np.loadtxt(f"{tag}_KerFeatures_{d}_NumProj_{P}_TopM_{M}_TopK_{K}", dtype=np.int64)
        oc, od, (ptr, idx, dd), _, _ = O.find_core(Xt, metric, tk, tm, float(e), minPts, mode=1)
        olab, onc = O.form_cluster(oc, ptr, idx)
        assert np.array_equal(lab, olab), f"sDbscan labels differ at eps {e}"
        n_clustered += int((olab >= 0).sum())

        slab = np.loadtxt(f"{tag}_Prob_{lround(np.float32(1000) * prob)}", dtype=np.int64)
        sc, sd, (sptr, sidx, sdd) = O.sng_find_core(Xt, metric, seq, float(e), minPts, win_end=win)
        solab, _ = O.form_cluster(sc, sptr, sidx)
        assert np.array_equal(slab, solab), f"sngDbscan labels differ at eps {e}"
    assert n_clustered > 0   # the sweep is not all noise


def test_cli_reports_negative_chi2_input(tmp_path):
    """transformData's error for a negative Chi2 entry, from the loader's
    transform, with exit code 1 (src/Utilities.h:308-312)."""
    X = np.abs(np.random.default_rng(1).standard_normal((300, 16))).astype(np.float32)
    X[17, 3] = -0.5
    data = tmp_path / "neg.txt"
    np.savetxt(data, X, fmt="%.9g")
    r = subprocess.run([CLI, "--n_points", "300", "--n_features", "16", "--eps", "0.3", "--minPts", "5",
                        "--distance", "Chi2", "--dataset", str(data)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Error: X is not non-negative !" in r.stderr


def test_transform_entry_point_matches_the_oracle(ctx):
    """sdb_transform (transformData, src/Utilities.h:289-335, as loadtxtData
    applies it): Cosine rows divided by their norm, Chi2 / JS rows by their
    sum, L2 rows unchanged -- bit-equal to the oracle's transform -- and the
    reference's errors for a negative entry and a zero point."""
    Xc, _ = datagen.mixture(3000, 100, seed=7)
    Xh, _ = datagen.dirichlet(3000, 100, k=5, seed=8)
    for metric, X in [(S.COSINE, Xc), (S.CHI2, Xh), (S.JS, Xh), (S.L2, Xc)]:
        print("sdb_transform metric", metric, flush=True)
        ctx.set_params(3000, 100, n_proj=256, topK=5, topM=20, metric=metric, ker_n_features=200)
        ctx.load(X)
        got = ctx.transform()
        want, rc = O.transform(X, metric)
        assert rc == 0
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), metric
    Xn = Xh.copy()
    Xn[17, 3] = -0.5
    ctx.set_params(3000, 100, n_proj=256, topK=5, topM=20, metric=S.CHI2, ker_n_features=200)
    ctx.load(Xn)
    with pytest.raises(S.SdbError) as e:
        ctx.transform()
    assert e.value.code == 4 and "not non-negative" in str(e.value)
    Xz = Xc.copy()
    Xz[5] = 0.0
    ctx.set_params(3000, 100, n_proj=256, topK=5, topM=20, metric=S.COSINE)
    ctx.load(Xz)
    with pytest.raises(S.SdbError) as e:
        ctx.transform()
    assert e.value.code == 5
