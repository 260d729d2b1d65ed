"""The command-line front end `sdbscan` (src/main.cpp:15-85 with
readParam_sDbscan, src/Utilities.cpp:109-468) on the CPU: argument errors,
the defaults the reader derives from the distance, and the missing-dataset
error.  None of these paths touches a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sdbscan_amd", "sdbscan")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="sdbscan CLI not built (run build())")


def run(*args):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


BASE = ["--n_points", 100, "--n_features", 51, "--eps", 0.5, "--minPts", 5]


def test_too_few_arguments():
    rc, out, err = run("--n_points", 100)
    assert rc == 1 and "Error: Not enough parameters !" in err


@pytest.mark.parametrize("drop,msg", [
    ("--n_points", "Error: Number of rows/points is missing !"),
    ("--n_features", "Error: Number of columns/dimensions is missing !"),
    ("--minPts", "Error: minPts is missing !"),
    ("--eps", "Error: Eps is missing !"),
])
def test_required_values(drop, msg):
    args = list(BASE)
    i = args.index(drop)
    del args[i:i + 2]
    rc, out, err = run(*args, "--verbose", "--topK", 3)   # keep at least 6 arguments
    assert rc == 1 and msg in err


def test_missing_dataset_after_parsing():
    rc, out, err = run(*BASE)
    assert rc == 1 and "Error: Data file does not exist !" in err
    rc, out, err = run(*BASE, "--dataset", "/nonexistent/file.txt")
    assert rc == 1 and "Error: Data file does not exist !" in err


def lines(out):
    return out.splitlines()


def test_cosine_defaults():
    rc, out, _ = run(*BASE)
    L = lines(out)
    assert "Distance is missing so we use default cosine distance" in L
    assert "TopK is missing. Use default topK: 5" in L
    assert "TopM is missing. Use default TopM = minPts = 5" in L
    assert "Kernel features is missing. Use default number of kernel features: 51" in L
    # max(256, 2^ceil(log2(n_features)))
    assert "Number of projections is missing. Use number of projections: 256" in L
    assert "Sampling probability is missing. Use default sampling ratio: 0.01" in L
    assert "No output file" in L
    assert "Number of threads is missing. Use all threads: -1" in L
    assert "Use a random seed: -1" in L
    assert "Default: We do not cluster the noisy points." in L
    assert not any(l.startswith("Sigma") for l in L)


@pytest.mark.parametrize("dist,ker,sigma", [("L1", 102, "0.5"), ("L2", 102, "1")])
def test_kernel_defaults_l1_l2(dist, ker, sigma):
    rc, out, _ = run(*BASE, "--distance", dist)
    L = lines(out)
    assert f"{dist} distance" in L
    assert f"Kernel features is missing. Use default number of kernel features: {ker}" in L
    what = "eps for L1" if dist == "L1" else "2 * eps for L2"
    assert f"Sigma is missing. Use default sigma = {what}: {sigma}" in L
    assert not any("Interval sampling ratio" in l for l in L)


def test_kernel_defaults_chi2_and_projections_from_kernel_features():
    rc, out, _ = run(*BASE, "--distance", "Chi2", "--ker_n_features", 600)
    L = lines(out)
    assert "Chi2 distance" in L
    assert "Number of kernel embedded dimensions: 600" in L
    # non-cosine distances size the projections by the kernel features
    assert "Number of projections is missing. Use number of projections: 1024" in L
    assert "Interval sampling ratio is missing. Use default sampling ratio for Chi2 and JS distances: 0.4" in L


def test_explicit_values_and_unknown_distance():
    rc, out, err = run(*BASE, "--distance", "Hamming", "--topK", 7, "--topM", 40, "--n_proj", 512,
                       "--clusterNoise", 1, "--samplingProb", 0.1, "--n_threads", 4, "--random_seed", 9,
                       "--output", "/tmp/x", "--verbose")
    L = lines(out)
    assert "Use default cosine distance" in L
    for want in ["TopK closest/furthest vectors: 7", "TopM: 40", "Number of projections: 512",
                 "We cluster noisy points.", "Sampling ratio for sngDbscan or sDbscan-1NN: 0.1",
                 "Number of threads: 4", "Seed: 9", "verbose = true."]:
        assert want in L, want
    assert "No output file" not in L
    assert rc == 1 and "Error: Data file does not exist !" in err
