"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle restates the reference's neighbourhood-discovery path on the CPU
(see oracle.cpp for the file:line each function follows).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker
or the timed baseline; the product (sdbscan_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-C", HERE, "liboracle.so"])
        L = ctypes.CDLL(LIB)
        sig = {
            "orc_hd3_bits": ([ctypes.c_int, ctypes.c_int, _P], None),
            "orc_gauss": ([ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P], None),
            "orc_cauchy": ([ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P], None),
            "orc_transform": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "orc_fht": ([_P, ctypes.c_int], None),
            "orc_label_noise": ([ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P], None),
            "orc_sng_find_core": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_int, _P, _P], _P),
            "orc_fwht_textbook": ([_P, ctypes.c_int], None),
            "orc_embed": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P, _P], None),
            "orc_project_embedded": ([_P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P], None),
            "orc_topk": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], None),
            "orc_heap_select": ([_P, ctypes.c_int, ctypes.c_int, _P, _P], None),
            "orc_topm": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], None),
            "orc_dist32": ([_P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_float),
            "orc_dist32_lanes": ([_P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_float),
            "orc_dist64": ([_P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_double),
            "orc_find_core": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P,
                               ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P],
                              ctypes.c_void_p),
            "orc_csr_size": ([_P], ctypes.c_longlong),
            "orc_csr_copy": ([_P, _P, _P, _P], None),
            "orc_free": ([_P], None),
            "orc_form_cluster": ([ctypes.c_int, _P, _P, _P, _P], ctypes.c_int),
            "orc_core_dist": ([ctypes.c_int, ctypes.c_int, _P, _P, _P, _P], None),
            "orc_optics": ([ctypes.c_int, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_fit_sdbscan": ([_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_float, _P, _P, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_int, ctypes.c_int, _P, _P, _P], ctypes.c_int),
            "orc_max_threads": ([], ctypes.c_int),
        }
        for k, (a, r) in sig.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def hd3_bits(nbits, seed):
    out = np.zeros(nbits, np.uint8)
    lib().orc_hd3_bits(nbits, seed, out.ctypes.data)
    return out


def gauss(rows, cols, mean, std, seed):
    out = np.zeros(rows * cols, np.float32)
    lib().orc_gauss(rows, cols, mean, std, seed, out.ctypes.data)
    return out


def cauchy(rows, cols, x0, gamma, seed):
    out = np.zeros(rows * cols, np.float32)
    lib().orc_cauchy(rows, cols, x0, gamma, seed, out.ctypes.data)
    return out


def transform(X, metric):
    X = _c(X, np.float32).copy()
    rc = lib().orc_transform(X.ctypes.data, X.shape[0], X.shape[1], metric)
    return X, rc


def fht(v, log_n):
    v = _c(v, np.float32).copy()
    lib().orc_fht(v.ctypes.data, log_n)
    return v


def embed(X, metric, ker, L, R=None):
    X = _c(X, np.float32)
    n, d = X.shape
    E = np.zeros((n, ker), np.float32)
    Rp = _c(R, np.float32).ctypes.data if R is not None else None
    Rk = _c(R, np.float32) if R is not None else None
    lib().orc_embed(X.ctypes.data, n, d, metric, ker, L, Rk.ctypes.data if Rk is not None else None,
                    E.ctypes.data)
    return E


def project(E, bits, D, P):
    E = _c(E, np.float32)
    n, ker = E.shape
    proj = np.zeros((P, n), np.float32)
    lib().orc_project_embedded(E.ctypes.data, n, ker, _c(bits, np.uint8).ctypes.data, D, P, proj.ctypes.data)
    return proj


def topk(proj, K):
    proj = _c(proj, np.float32)
    P, n = proj.shape
    out = np.zeros((n, 2 * K), np.int32)
    lib().orc_topk(proj.ctypes.data, n, P, K, out.ctypes.data)
    return out


def topm(proj, M):
    proj = _c(proj, np.float32)
    P, n = proj.shape
    out = np.zeros((P, 2 * M), np.int32)
    lib().orc_topm(proj.ctypes.data, n, P, M, out.ctypes.data)
    return out


def heap_select(vals, k):
    vals = _c(vals, np.float32)
    c = np.zeros(k, np.int32)
    f = np.zeros(k, np.int32)
    lib().orc_heap_select(vals.ctypes.data, len(vals), k, c.ctypes.data, f.ctypes.data)
    return c, f


def dist32(x, y, metric):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    return lib().orc_dist32(x.ctypes.data, y.ctypes.data, len(x), metric)


def dist32_lanes(x, y, metric):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    return lib().orc_dist32_lanes(x.ctypes.data, y.ctypes.data, len(x), metric)


def dist64(x, y, metric):
    x = _c(x, np.float32)
    y = _c(y, np.float32)
    return lib().orc_dist64(x.ctypes.data, y.ctypes.data, len(x), metric)


def find_core(X, metric, topK, topM, eps, minPts, mode=1):
    """Returns core (n,), degree (n,), CSR (ptr, idx, dist32), near_pairs, n_dist.
    mode bit 0: decide in FP64 (else sequential FP32); bit 1: CSR distances in
    the device's lane order (dist32_lanes)."""
    X = _c(X, np.float32)
    n, d = X.shape
    topK = _c(topK, np.int32)
    topM = _c(topM, np.int32)
    K = topK.shape[1] // 2
    M = topM.shape[1] // 2
    core = np.zeros(n, np.uint8)
    deg = np.zeros(n, np.int32)
    near = ctypes.c_longlong()
    nd = ctypes.c_longlong()
    h = lib().orc_find_core(X.ctypes.data, n, d, metric, topK.ctypes.data, K, topM.ctypes.data, M,
                            eps, minPts, mode, core.ctypes.data, deg.ctypes.data,
                            ctypes.byref(near), ctypes.byref(nd))
    tot = lib().orc_csr_size(h)
    ptr = np.zeros(n + 1, np.int32)
    idx = np.zeros(max(tot, 1), np.int32)
    dd = np.zeros(max(tot, 1), np.float32)
    lib().orc_csr_copy(h, ptr.ctypes.data, idx.ctypes.data, dd.ctypes.data)
    lib().orc_free(h)
    return core, deg, (ptr, idx[:tot], dd[:tot]), near.value, nd.value


def sng_find_core(X, metric, samples, eps, minPts, win_end=None, both_ways=True, dedup=False):
    """sngDbscan / sngOptics neighbourhoods from a draw sequence: core, degree,
    deduplicated CSR (ptr, idx, dist32)."""
    X = _c(X, np.float32)
    n, d = X.shape
    smp = _c(samples, np.int32)
    we = _c(win_end, np.int32) if win_end is not None else None
    core = np.zeros(n, np.uint8)
    deg = np.zeros(n, np.int32)
    flags = (1 if both_ways else 0) | (2 if dedup else 0)
    h = lib().orc_sng_find_core(X.ctypes.data, n, d, metric, smp.ctypes.data if len(smp) else None, len(smp),
                                we.ctypes.data if we is not None else None, flags, eps, minPts,
                                core.ctypes.data, deg.ctypes.data)
    tot = lib().orc_csr_size(h)
    ptr = np.zeros(n + 1, np.int32)
    idx = np.zeros(max(tot, 1), np.int32)
    dd = np.zeros(max(tot, 1), np.float32)
    lib().orc_csr_copy(h, ptr.ctypes.data, idx.ctypes.data, dd.ctypes.data)
    lib().orc_free(h)
    return core, deg, (ptr, idx[:tot], dd[:tot])


def form_cluster(core, ptr, idx):
    core = _c(core, np.uint8)
    ptr = _c(ptr, np.int32)
    idx = _c(idx, np.int32)
    n = len(core)
    lab = np.zeros(n, np.int32)
    nc = lib().orc_form_cluster(n, core.ctypes.data, ptr.ctypes.data,
                                idx.ctypes.data if len(idx) else np.zeros(1, np.int32).ctypes.data,
                                lab.ctypes.data)
    return lab, nc


def label_noise(cproj, topK, labels, sampled):
    """sDbscan-1NN labels (src/sDbscan.cpp:775-806); cproj: P x nS projections
    of the sampled core points, in the order of `sampled`."""
    cproj = _c(cproj, np.float32)
    topK = _c(topK, np.int32)
    sampled = _c(sampled, np.int32)
    lab = np.array(labels, np.int32, copy=True)
    P, nS = cproj.shape
    n, K2 = topK.shape
    lib().orc_label_noise(P, nS, cproj.ctypes.data, n, K2 // 2, topK.ctypes.data, sampled.ctypes.data,
                          lab.ctypes.data)
    return lab


def core_dist(minPts, ptr, dd):
    ptr = _c(ptr, np.int32)
    dd = _c(dd, np.float32) if len(dd) else np.zeros(1, np.float32)
    n = len(ptr) - 1
    core = np.zeros(n, np.uint8)
    cd = np.zeros(n, np.float32)
    lib().orc_core_dist(n, minPts, ptr.ctypes.data, dd.ctypes.data, core.ctypes.data, cd.ctypes.data)
    return core, cd


def optics(core, cd, ptr, idx, dd):
    n = len(core)
    ptr = _c(ptr, np.int32)
    idx = _c(idx, np.int32) if len(idx) else np.zeros(1, np.int32)
    dd = _c(dd, np.float32) if len(dd) else np.zeros(1, np.float32)
    order = np.zeros(n, np.int32)
    reach = np.zeros(n, np.float32)
    lib().orc_optics(n, _c(core, np.uint8).ctypes.data, _c(cd, np.float32).ctypes.data, ptr.ctypes.data,
                     idx.ctypes.data, dd.ctypes.data, order.ctypes.data, reach.ctypes.data)
    return order, reach


def fit_sdbscan(X, metric, P, K, M, ker, L, R, bits, D, eps, minPts, n_threads=0):
    """Full CPU fit with the reference's phase timers; X is transformed in place."""
    X = _c(X, np.float32).copy()
    n, d = X.shape
    lab = np.zeros(n, np.int32)
    core = np.zeros(n, np.uint8)
    t = np.zeros(4, np.float64)
    Rk = _c(R, np.float32) if R is not None else None
    nc = lib().orc_fit_sdbscan(X.ctypes.data, n, d, metric, P, K, M, ker, L,
                               Rk.ctypes.data if Rk is not None else None,
                               _c(bits, np.uint8).ctypes.data, D, eps, minPts, n_threads,
                               lab.ctypes.data, core.ctypes.data, t.ctypes.data)
    return lab, nc, core, t


def max_threads():
    return lib().orc_max_threads()
