"""sdbscan_amd -- sDbscan's neighbourhood-discovery path on AMD Instinct MI355X.

The product is two in-tree shared libraries built by ``sdbscan_amd.build``:

* ``libsdbscan_amd.so`` -- HIP kernels for gfx950 behind the C ABI declared in
  ``include/sdbscan_amd.h``;
* ``sDbscan.cpython-*.so`` -- the pybind11 module ``sDbscan``, a drop-in for
  the reference's module (same class, methods and keyword arguments).

``capi()`` returns the C ABI through ctypes; ``module()`` imports the drop-in
module.  Both raise if the libraries are not built: there is no CPU fallback.
"""
import ctypes
import glob
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdbscan_amd.so")

COSINE, L1, L2, CHI2, JS = 0, 1, 2, 3, 4
FP32, BF16 = 0, 1
NOALLOC_MAX_SAMPLES = 1 << 18  # SDB_NOALLOC_MAX_SAMPLES
METRICS = {"Cosine": COSINE, "L1": L1, "L2": L2, "Chi2": CHI2, "JS": JS}


class Params(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_longlong), ("n_features", ctypes.c_int),
                ("n_proj", ctypes.c_int), ("topK", ctypes.c_int), ("topM", ctypes.c_int),
                ("metric", ctypes.c_int), ("ker_n_features", ctypes.c_int),
                ("ker_sigma", ctypes.c_float), ("ker_intervalSampling", ctypes.c_float),
                ("fhtDim", ctypes.c_int), ("precision", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("t_transform_ms", ctypes.c_double), ("t_index_ms", ctypes.c_double),
                ("t_core_ms", ctypes.c_double), ("t_cluster_ms", ctypes.c_double),
                ("t_dist_ms", ctypes.c_double), ("n_pairs", ctypes.c_longlong),
                ("n_queue", ctypes.c_longlong), ("n_recheck32", ctypes.c_longlong),
                ("n_recheck64", ctypes.c_longlong), ("n_recheck_hits", ctypes.c_longlong),
                ("n_hubs", ctypes.c_longlong), ("n_tie_points", ctypes.c_longlong),
                ("n_tie_rows", ctypes.c_longlong), ("n_core", ctypes.c_longlong),
                ("n_components", ctypes.c_int), ("band", ctypes.c_double),
                ("band32", ctypes.c_double), ("n_near_eps", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchJob(ctypes.Structure):
    """sdb_batch_job of include/sdbscan_amd.h."""
    _fields_ = [("X", ctypes.c_void_p), ("params", Params), ("hd3_bits", ctypes.c_void_p),
                ("R", ctypes.c_void_p), ("n_eps", ctypes.c_int), ("eps", ctypes.c_void_p),
                ("n_minpts", ctypes.c_int), ("minPts", ctypes.c_void_p), ("optics", ctypes.c_int),
                ("labels", ctypes.c_void_p), ("n_clusters", ctypes.c_void_p),
                ("ordering", ctypes.c_void_p), ("reach", ctypes.c_void_p), ("status", ctypes.c_int),
                ("tile_rows", ctypes.c_int), ("ms", ctypes.c_double), ("error", ctypes.c_char * 192)]


_P = ctypes.c_void_p
_SIGS = {
    "sdb_hd3_bits": ([ctypes.c_int, ctypes.c_uint, _P], ctypes.c_int),
    "sdb_rff_matrix": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, _P],
                       ctypes.c_int),
    "sdb_create": ([ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "sdb_destroy": ([_P], None),
    "sdb_last_error": ([_P], ctypes.c_char_p),
    "sdb_set_params": ([_P, ctypes.POINTER(Params)], ctypes.c_int),
    "sdb_load_host": ([_P, _P, ctypes.c_longlong, ctypes.c_int], ctypes.c_int),
    "sdb_load_device": ([_P, _P, ctypes.c_longlong, ctypes.c_int], ctypes.c_int),
    "sdb_build_index": ([_P, _P, _P], ctypes.c_int),
    "sdb_find_core": ([_P, ctypes.c_float, ctypes.c_int, _P, _P], ctypes.c_int),
    "sdb_form_clusters": ([_P, _P, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "sdb_label_noise": ([_P, _P, ctypes.c_int, _P], ctypes.c_int),
    "sdb_sng_find_core": ([_P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_float, ctypes.c_int, _P, _P],
                          ctypes.c_int),
    "sdb_get_index": ([_P, _P, _P], ctypes.c_int),
    "sdb_get_projections": ([_P, _P, ctypes.c_int, _P], ctypes.c_int),
    "sdb_get_transformed": ([_P, _P], ctypes.c_int),
    "sdb_transform": ([_P, _P], ctypes.c_int),
    "sdb_get_transformed_bf16": ([_P, _P], ctypes.c_int),
    "sdb_get_neighbors": ([_P, _P, _P, _P], ctypes.c_int),
    "sdb_core_dist": ([_P, ctypes.c_int, _P], ctypes.c_int),
    "sdb_optics_reach": ([_P, _P, _P], ctypes.c_int),
    "sdb_optics_order": ([_P, ctypes.c_int, _P, _P], ctypes.c_int),
    "sdb_optics_orders": ([_P, ctypes.c_int, _P, _P, _P], ctypes.c_int),
    "sdb_set_tile_rows": ([_P, ctypes.c_int], ctypes.c_int),
    "sdb_tune_tile_rows": ([_P, ctypes.c_float, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "sdb_batch_sweep": ([ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "sdb_batch_run": ([ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "sdb_get_stats": ([_P, ctypes.POINTER(Stats)], ctypes.c_int),
    "sdb_sample_pair_dist": ([_P, ctypes.c_longlong, ctypes.c_ulonglong, _P, _P, _P, _P], ctypes.c_int),
    "sdb_sample_pair_dist_noalloc": ([_P, ctypes.c_longlong, ctypes.c_ulonglong, _P, _P, _P, _P], ctypes.c_int),
    "sdb_generate": ([_P, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_ulonglong,
                      ctypes.c_float], ctypes.c_int),
    "sdb_copy_input": ([_P, ctypes.c_longlong, ctypes.c_longlong, _P], ctypes.c_int),
}

_lib = None


def capi():
    """The C ABI of libsdbscan_amd.so (ctypes, signatures set)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m sdbscan_amd.build")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def module():
    """The pybind11 drop-in module ``sDbscan``."""
    cands = glob.glob(os.path.join(HERE, "sDbscan*.so"))
    if not cands:
        raise RuntimeError("sDbscan module is not built: run python -m sdbscan_amd.build")
    spec = importlib.util.spec_from_file_location("sDbscan", cands[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def fht_dim(n_proj, ker_n_features):
    """fhtDim of src/sDbscan.h:107."""
    import math
    return 1 << int(math.ceil(math.log2(max(n_proj, ker_n_features))))


def hd3_bits(n_bits, seed):
    """The 3 * fhtDim HD3 sign bytes of bitHD3Generator (host std:: calls)."""
    import numpy as np
    out = np.zeros(n_bits, np.uint8)
    rc = capi().sdb_hd3_bits(n_bits, seed, out.ctypes.data)
    if rc != 0:
        raise SdbError(rc, "sdb_hd3_bits failed")
    return out


def rff_matrix(metric, half, d, sigma, seed):
    """The (half x d column-major) random matrix of the L1 / L2 embeddings."""
    import numpy as np
    out = np.zeros(half * d, np.float32)
    rc = capi().sdb_rff_matrix(metric, half, d, sigma, seed, out.ctypes.data)
    if rc != 0:
        raise SdbError(rc, "sdb_rff_matrix failed")
    return out


class SdbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Context:
    """Thin ctypes wrapper over one sdb_ctx (tests and bench use it)."""

    def __init__(self, device=0):
        self.lib = capi()
        h = ctypes.c_void_p()
        rc = self.lib.sdb_create(device, ctypes.byref(h))
        if rc != 0:
            raise SdbError(rc, "sdb_create failed (no HIP device?)")
        self.h = h

    def close(self):
        if self.h:
            self.lib.sdb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise SdbError(rc, self.lib.sdb_last_error(self.h).decode())

    def set_params(self, n, d, n_proj=1024, topK=10, topM=50, metric=COSINE, ker_n_features=None,
                   ker_sigma=1.0, ker_intervalSampling=0.4, precision=FP32):
        if ker_n_features is None or metric == COSINE:
            ker_n_features = d if metric == COSINE else 1024
        p = Params(n, d, n_proj, topK, topM, metric, ker_n_features, ker_sigma,
                   ker_intervalSampling, fht_dim(n_proj, ker_n_features), precision)
        self.params = p
        self._chk(self.lib.sdb_set_params(self.h, ctypes.byref(p)))
        return p

    def load(self, X):
        import numpy as np
        X = np.ascontiguousarray(X, dtype=np.float32)
        self._chk(self.lib.sdb_load_host(self.h, X.ctypes.data, X.shape[0], X.shape[1]))

    def generate(self, n, d, seed, sigma=0.1, kind=0):
        self._chk(self.lib.sdb_generate(self.h, kind, n, d, seed, sigma))

    def build_index(self, bits, R=None):
        import numpy as np
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        Rp = None
        if R is not None:
            R = np.ascontiguousarray(R, dtype=np.float32)
            Rp = R.ctypes.data
        self._chk(self.lib.sdb_build_index(self.h, bits.ctypes.data, Rp))

    def find_core(self, eps, minPts, want=True):
        import numpy as np
        n = self.params.n_points
        core = np.zeros(n, np.uint8) if want else None
        deg = np.zeros(n, np.int32) if want else None
        self._chk(self.lib.sdb_find_core(self.h, eps, minPts,
                                         core.ctypes.data if want else None,
                                         deg.ctypes.data if want else None))
        return core, deg

    def form_clusters(self, want=True):
        import numpy as np
        n = self.params.n_points
        lab = np.zeros(n, np.int32) if want else None
        nc = ctypes.c_int()
        self._chk(self.lib.sdb_form_clusters(self.h, lab.ctypes.data if want else None,
                                             ctypes.byref(nc)))
        return lab, nc.value

    def sng_find_core(self, samples, eps, minPts, win_end=None, both_ways=True, dedup=False):
        """Neighbourhoods from a sample sequence (sngDbscan / sngOptics):
        core flags and degrees."""
        import numpy as np
        n = self.params.n_points
        smp = np.ascontiguousarray(samples, dtype=np.int32)
        we = None if win_end is None else np.ascontiguousarray(win_end, dtype=np.int32)
        core = np.zeros(n, np.uint8)
        deg = np.zeros(n, np.int32)
        flags = (1 if both_ways else 0) | (2 if dedup else 0)
        self._chk(self.lib.sdb_sng_find_core(self.h, smp.ctypes.data if len(smp) else None, len(smp),
                                             we.ctypes.data if we is not None else None, flags,
                                             eps, minPts, core.ctypes.data, deg.ctypes.data))
        return core, deg

    def label_noise(self, sampled_core):
        """sDbscan-1NN labels of the noise points after form_clusters, for the
        given sampled core points (increasing index order)."""
        import numpy as np
        ids = np.ascontiguousarray(sampled_core, dtype=np.int32)
        lab = np.zeros(self.params.n_points, np.int32)
        self._chk(self.lib.sdb_label_noise(self.h, ids.ctypes.data if len(ids) else None, len(ids),
                                           lab.ctypes.data))
        return lab

    def get_index(self):
        import numpy as np
        p = self.params
        tk = np.zeros((p.n_points, 2 * p.topK), np.int32)
        tm = np.zeros((p.n_proj, 2 * p.topM), np.int32)
        self._chk(self.lib.sdb_get_index(self.h, tk.ctypes.data, tm.ctypes.data))
        return tk, tm

    def get_projections(self, ids):
        """HD3 projections of the listed points: (len(ids), n_proj) float32,
        row q = the MATRIX_FHT column of point ids[q]."""
        import numpy as np
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.zeros((len(ids), self.params.n_proj), np.float32)
        self._chk(self.lib.sdb_get_projections(self.h, ids.ctypes.data, len(ids), out.ctypes.data))
        return out

    def get_transformed(self):
        import numpy as np
        p = self.params
        X = np.zeros((p.n_points, p.n_features), np.float32)
        self._chk(self.lib.sdb_get_transformed(self.h, X.ctypes.data))
        return X

    def transform(self):
        """transformData (src/Utilities.h:289-335) of the loaded input; returns
        the transformed n x d rows."""
        import numpy as np
        p = self.params
        X = np.zeros((p.n_points, p.n_features), np.float32)
        self._chk(self.lib.sdb_transform(self.h, X.ctypes.data))
        return X

    def get_transformed_bf16(self):
        """BF16 bit patterns (uint16) of the transformed points."""
        import numpy as np
        p = self.params
        X = np.zeros((p.n_points, p.n_features), np.uint16)
        self._chk(self.lib.sdb_get_transformed_bf16(self.h, X.ctypes.data))
        return X

    def get_neighbors(self):
        import numpy as np
        n = self.params.n_points
        ptr = np.zeros(n + 1, np.int64)
        self._chk(self.lib.sdb_get_neighbors(self.h, ptr.ctypes.data, None, None))
        tot = int(ptr[-1])
        idx = np.zeros(max(tot, 1), np.int32)
        dist = np.zeros(max(tot, 1), np.float32)
        self._chk(self.lib.sdb_get_neighbors(self.h, ptr.ctypes.data, idx.ctypes.data,
                                             dist.ctypes.data))
        return ptr, idx[:tot], dist[:tot]

    def core_dist(self, minPts):
        import numpy as np
        out = np.zeros(self.params.n_points, np.float32)
        self._chk(self.lib.sdb_core_dist(self.h, minPts, out.ctypes.data))
        return out

    def optics_reach(self, ordering):
        import numpy as np
        ordering = np.ascontiguousarray(ordering, dtype=np.int32)
        out = np.zeros(self.params.n_points, np.float32)
        self._chk(self.lib.sdb_optics_reach(self.h, ordering.ctypes.data, out.ctypes.data))
        return out

    def optics_order(self, minPts):
        """sOPTICS ordering and reachability at the last find_core's eps."""
        import numpy as np
        n = self.params.n_points
        order = np.zeros(n, np.int32)
        reach = np.zeros(n, np.float32)
        self._chk(self.lib.sdb_optics_order(self.h, minPts, order.ctypes.data, reach.ctypes.data))
        return order, reach

    def optics_orders(self, minPts):
        """optics_order for several minPts at once: arrays [len(minPts), n]."""
        import numpy as np
        n = self.params.n_points
        mp = np.ascontiguousarray(minPts, dtype=np.int32)
        order = np.zeros((len(mp), n), np.int32)
        reach = np.zeros((len(mp), n), np.float32)
        self._chk(self.lib.sdb_optics_orders(self.h, len(mp), mp.ctypes.data, order.ctypes.data,
                                             reach.ctypes.data))
        return order, reach

    def set_tile_rows(self, rows):
        self._chk(self.lib.sdb_set_tile_rows(self.h, rows))

    def tune_tile_rows(self, eps, candidates):
        import numpy as np
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        best = ctypes.c_int()
        ms = ctypes.c_double()
        self._chk(self.lib.sdb_tune_tile_rows(self.h, eps, cand.ctypes.data, len(cand), ctypes.byref(best),
                                              ctypes.byref(ms)))
        return best.value, ms.value

    def stats(self):
        s = Stats()
        self._chk(self.lib.sdb_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def sample_pair_dist(self, n_samples, seed=1):
        """Candidate pairs drawn from the index with the tiles' value of the
        tested quantity and its FP64 value: (a, b, fast, exact)."""
        import numpy as np
        a = np.zeros(n_samples, np.int32)
        b = np.zeros(n_samples, np.int32)
        f = np.zeros(n_samples, np.float32)
        e = np.zeros(n_samples, np.float64)
        self._chk(self.lib.sdb_sample_pair_dist(self.h, n_samples, seed, a.ctypes.data, b.ctypes.data,
                                                 f.ctypes.data, e.ctypes.data))
        return a, b, f, e

    def sample_pair_dist_noalloc(self, n_samples, seed=1):
        """sample_pair_dist through the entry point that allocates no device
        memory (at most NOALLOC_MAX_SAMPLES pairs)."""
        import numpy as np
        a = np.zeros(n_samples, np.int32)
        b = np.zeros(n_samples, np.int32)
        f = np.zeros(n_samples, np.float32)
        e = np.zeros(n_samples, np.float64)
        self._chk(self.lib.sdb_sample_pair_dist_noalloc(self.h, n_samples, seed, a.ctypes.data, b.ctypes.data,
                                                         f.ctypes.data, e.ctypes.data))
        return a, b, f, e

    def copy_input(self, row0, rows):
        import numpy as np
        out = np.zeros((rows, self.params.n_features), np.float32)
        self._chk(self.lib.sdb_copy_input(self.h, row0, rows, out.ctypes.data))
        return out


def _batch_jobs(datasets, eps, minPts, optics, n_proj, topK, topM, metric, seed, precision, hd3):
    import numpy as np
    if hd3 is None:
        hd3 = hd3_bits
    if metric in (L1, L2):
        raise ValueError("batch runs build Cosine / Chi2 / JS jobs; L1 / L2 jobs need an RFF matrix per job")
    eps_a = np.ascontiguousarray(eps, dtype=np.float32)
    mp_a = np.ascontiguousarray(minPts, dtype=np.int32)
    jobs = (BatchJob * len(datasets))()
    keep = [eps_a, mp_a]
    out = []
    for i, X in enumerate(datasets):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d = X.shape
        ker = d if metric == COSINE else 1024
        D = fht_dim(n_proj, ker)
        bits = np.ascontiguousarray(hd3(3 * D, seed), dtype=np.uint8)
        lab = np.zeros((len(eps_a), len(mp_a), n), np.int32)
        ncl = np.zeros((len(eps_a), len(mp_a)), np.int32)
        order = np.zeros((len(mp_a), n), np.int32)
        reach = np.zeros((len(mp_a), n), np.float32)
        keep += [X, bits, lab, ncl, order, reach]
        j = jobs[i]
        j.X = X.ctypes.data
        j.params = Params(n, d, n_proj, topK, topM, metric, ker, 1.0, 0.4, D, precision)
        j.hd3_bits = bits.ctypes.data
        j.R = None
        j.n_eps, j.eps = len(eps_a), eps_a.ctypes.data
        j.n_minpts, j.minPts = len(mp_a), mp_a.ctypes.data
        j.optics = 1 if optics else 0
        j.labels, j.n_clusters = lab.ctypes.data, ncl.ctypes.data
        j.ordering, j.reach = order.ctypes.data, reach.ctypes.data
        out.append({"labels": lab, "n_clusters": ncl, "ordering": order, "reach": reach})
    return jobs, keep, out


def _batch_results(jobs, out, rc):
    for i in range(len(out)):
        out[i]["status"] = jobs[i].status
        out[i]["error"] = jobs[i].error.decode()
        out[i]["tile_rows"] = jobs[i].tile_rows
        out[i]["ms"] = jobs[i].ms
    if rc != 0:
        bad = [o["error"] for o in out if o["status"] != 0]
        raise SdbError(rc, "; ".join(bad[:3]))
    return out


def batch_run(datasets, eps, minPts, optics=True, n_streams=4, autotune=True, device=0, n_proj=1024, topK=10,
              topM=50, metric=COSINE, seed=42, precision=FP32, hd3=None):
    """Run sdb_batch_run over a list of (n x d) float32 arrays (Cosine,
    Chi2 or JS), every dataset swept over the same eps and minPts values.
    ``hd3(n_bits, seed)`` gives the HD3 sign bytes (default: hd3_bits, the
    reference's bitHD3Generator).  Returns (results, wall_ms) with one dict
    per dataset:
    labels [n_eps, n_minpts, n], n_clusters [n_eps, n_minpts], ordering and
    reach [n_minpts, n] (optics), tile_rows and ms."""
    jobs, keep, out = _batch_jobs(datasets, eps, minPts, optics, n_proj, topK, topM, metric, seed, precision, hd3)
    wall = ctypes.c_double()
    rc = capi().sdb_batch_run(device, ctypes.cast(jobs, ctypes.c_void_p), len(datasets), n_streams,
                              1 if autotune else 0, ctypes.byref(wall))
    return _batch_results(jobs, out, rc), wall.value


SWEEP_STAGES = ("setup", "index", "autotune", "tiles", "recheck", "cluster_buffers", "capture", "chain", "copy_out",
                "optics", "teardown")


def batch_sweep(datasets, eps, minPts, optics=True, n_streams=4, autotune=True, device=0, n_proj=1024, topK=10,
                topM=50, seed=42, hd3=None):
    """The same sweep through sdb_batch_sweep: the candidate distances of all
    datasets in one persistent launch deciding every eps, the cluster chains
    of all (dataset, eps, minPts) sets one launch per step, captured as one
    graph (Cosine, FP32).  Returns (results, wall_ms, stages)
    with stages a dict of per-stage milliseconds (SWEEP_STAGES)."""
    jobs, keep, out = _batch_jobs(datasets, eps, minPts, optics, n_proj, topK, topM, COSINE, seed, FP32, hd3)
    wall = ctypes.c_double()
    stages = (ctypes.c_double * len(SWEEP_STAGES))()
    rc = capi().sdb_batch_sweep(device, ctypes.cast(jobs, ctypes.c_void_p), len(datasets), n_streams,
                                1 if autotune else 0, stages, ctypes.byref(wall))
    return _batch_results(jobs, out, rc), wall.value, dict(zip(SWEEP_STAGES, list(stages)))
