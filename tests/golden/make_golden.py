"""Writes the golden fixtures under tests/golden/ (run in the build container).

* fwht_golden.npz -- input/output vectors of the reference FFHT routine
  ``fht_float`` (src/fht.h:11, src/fht_avx.c:10081-10206) compiled from the
  reference sources by oracle/Makefile into oracle/_ref/libreffht.so, for
  log_n = 1..13, plus full HD3 chains (3 x {sign flip, fht_float(log_n=10)},
  src/sDbscan.cpp:74-91) on zero-padded cosine points.
* rng_kat.npz -- the first outputs of the host RNG calls of
  src/Utilities.h:47-122 (default_random_engine + uniform_int / normal /
  cauchy) made by the oracle's std:: calls; a libstdc++ drift check.

Usage:  make -C oracle && python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def load_ref():
    path = os.path.join(ROOT, "oracle", "_ref", "libreffht.so")
    lib = ctypes.CDLL(path)
    lib.fht_float.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.fht_float.restype = ctypes.c_int
    return lib


def load_oracle():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    lib.orc_hd3_bits.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.orc_gauss.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                              ctypes.c_int, ctypes.c_void_p]
    lib.orc_cauchy.argtypes = lib.orc_gauss.argtypes
    return lib


def fht(lib, v, log_n):
    buf = np.ascontiguousarray(v, dtype=np.float32).copy()
    rc = lib.fht_float(buf.ctypes.data, log_n)
    assert rc == 0
    return buf


def input_patterns(rng, n):
    """Inputs that exercise rounding, exact ties and signed zeros."""
    pats = [rng.standard_normal(n).astype(np.float32),
            rng.integers(-3, 4, n).astype(np.float32),
            (rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)]
    z = rng.standard_normal(n).astype(np.float32)
    z[n // 2:] = 0.0
    z[rng.random(n) < 0.25] = -0.0
    pats.append(z)
    # all signed zeros: the only inputs whose outputs tell FFHT's (0 - v)
    # butterflies apart from the textbook u + v, u - v
    sz = np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(0.0)).astype(np.float32)
    pats.append(sz)
    return pats


def main():
    ref = load_ref()
    orc = load_oracle()
    rng = np.random.default_rng(20240522)
    out = {}
    for log_n in range(1, 14):
        n = 1 << log_n
        pats = input_patterns(rng, n)
        if log_n >= 11:
            pats = pats[:2] + pats[3:]
        xin = np.stack(pats)
        xout = np.stack([fht(ref, v, log_n) for v in xin])
        out[f"in_{log_n}"] = xin
        out[f"out_{log_n}"] = xout

    # HD3 chains on zero-padded unit vectors (d = 784 -> D = 1024).
    D, d = 1024, 784
    bits = np.zeros(3 * D, dtype=np.uint8)
    orc.orc_hd3_bits(3 * D, 42, bits.ctypes.data)
    pts = np.abs(rng.standard_normal((6, d))).astype(np.float32)
    pts[:, rng.random(d) < 0.8] = 0.0
    pts /= np.linalg.norm(pts, axis=1, keepdims=True).astype(np.float32)
    chains = []
    for p in pts:
        v = np.zeros(D, dtype=np.float32)
        v[:d] = p
        for r in range(3):
            v = v * (2.0 * bits[r * D:(r + 1) * D].astype(np.float32) - 1.0).astype(np.float32)
            v = fht(ref, v, 10)
        chains.append(v)
    out["hd3_bits_seed42"] = bits
    out["hd3_in"] = pts
    out["hd3_out"] = np.stack(chains)
    np.savez_compressed(os.path.join(HERE, "fwht_golden.npz"), **out)

    kat = {}
    for seed in (0, 42):
        b = np.zeros(4096, dtype=np.uint8)
        orc.orc_hd3_bits(4096, seed, b.ctypes.data)
        kat[f"bits_{seed}"] = b
        g = np.zeros(1024, dtype=np.float32)
        orc.orc_gauss(256, 4, 0.0, 0.5, seed, g.ctypes.data)
        kat[f"gauss_{seed}"] = g
        c = np.zeros(1024, dtype=np.float32)
        orc.orc_cauchy(256, 4, 0.0, 0.5, seed, c.ctypes.data)
        kat[f"cauchy_{seed}"] = c
    np.savez_compressed(os.path.join(HERE, "rng_kat.npz"), **kat)
    print("wrote fwht_golden.npz and rng_kat.npz")


if __name__ == "__main__":
    sys.exit(main())
