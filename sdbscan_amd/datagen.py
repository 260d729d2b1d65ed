"""Seeded synthetic workloads (numpy, host side) for tests and the CPU baseline.

mixture()   -- 10-component sparse nonnegative mixture, L2-normalised
               (the 20k x 784 cosine workload and its 8M x 784 scale-up).
blobs()     -- anisotropic Gaussian blobs plus uniform noise (L2 workload).
dirichlet() -- Dirichlet histograms with cluster-specific boosted bins (Chi2).
sphere_mixture() -- Gaussian mixture normalised to the unit sphere (the
               many-dataset batch sweep).
"""
import numpy as np


def mixture(n, d=784, k=10, sigma=0.1, seed=1000, zero_frac=0.8):
    rng = np.random.default_rng(seed)
    centers = rng.random((k, d)).astype(np.float32)
    centers[rng.random((k, d)) < zero_frac] = 0.0
    comp = rng.integers(0, k, n)
    X = centers[comp] + rng.normal(0.0, sigma, (n, d)).astype(np.float32)
    np.clip(X, 0.0, 1.0, out=X)
    nrm = np.linalg.norm(X, axis=1, keepdims=True)
    nrm[nrm == 0] = 1.0
    return (X / nrm).astype(np.float32), comp


def blobs(n, d=51, k=12, noise=0.05, seed=1001, max_cond=10.0):
    rng = np.random.default_rng(seed)
    n_noise = int(n * noise)
    n_in = n - n_noise
    comp = rng.integers(0, k, n_in)
    centers = rng.uniform(-10, 10, (k, d))
    X = np.empty((n, d), np.float32)
    for c in range(k):
        q, _ = np.linalg.qr(rng.normal(size=(d, d)))
        ev = np.exp(rng.uniform(0, np.log(max_cond), d)) * 0.05
        A = q * np.sqrt(ev)
        idx = np.where(comp == c)[0]
        X[idx] = (centers[c] + rng.normal(size=(len(idx), d)) @ A.T).astype(np.float32)
    lo, hi = X[:n_in].min(0), X[:n_in].max(0)
    X[n_in:] = rng.uniform(lo, hi, (n_noise, d)).astype(np.float32)
    labels = np.concatenate([comp, -np.ones(n_noise, np.int64)])
    return X, labels


def dirichlet(n, d=784, k=10, base=0.05, boost=2.0, n_boost=20, seed=1004):
    rng = np.random.default_rng(seed)
    alphas = np.full((k, d), base)
    for c in range(k):
        alphas[c, rng.choice(d, n_boost, replace=False)] = boost
    comp = rng.integers(0, k, n)
    X = np.empty((n, d), np.float32)
    for c in range(k):
        idx = np.where(comp == c)[0]
        X[idx] = rng.dirichlet(alphas[c], len(idx)).astype(np.float32)
    return X, comp


def sphere_mixture(n, d, k, seed, spread=0.35):
    """k Gaussian components around random unit centres, each point
    normalised to the unit sphere (float32, n x d)."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(size=(k, d))
    centers /= np.linalg.norm(centers, axis=1, keepdims=True)
    comp = rng.integers(0, k, n)
    X = rng.standard_normal((n, d), dtype=np.float32)
    X *= np.float32(spread / np.sqrt(d))
    X += centers.astype(np.float32)[comp]
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return X, comp


def median_pairwise_l2(X, n_sample=10000, seed=1001):
    """Median L2 distance over the pairs of a seeded sample of rows (the RFF
    bandwidth sigma of the 1.8M x 51 L2 workload)."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    smp = X[rng.choice(n, min(n_sample, n), replace=False)].astype(np.float64)
    sq = (smp * smp).sum(1)
    D2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * smp @ smp.T, 0.0)
    return float(np.sqrt(np.median(D2[np.triu_indices(len(smp), 1)])))


def l2_eps_for_embedded_cosine(sigma, eps_cos):
    """Input-space L2 radius whose embedded cosine distance is eps_cos.  The
    normalised random Fourier features of N(0, 1/sigma) draws approximate the
    Gaussian kernel, so 1 - <e(x), e(y)> ~ 1 - exp(-r^2 / (2 sigma^2)) for
    ||x - y|| = r; eps_cos maps to r = sigma sqrt(-2 ln(1 - eps_cos)).  The
    reference decides L2 on the input rows (src/Utilities.h:218-219)."""
    return float(sigma * np.sqrt(-2.0 * np.log(1.0 - eps_cos)))
