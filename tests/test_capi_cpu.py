"""The C ABI library and the drop-in module load and export their interface
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import sdbscan_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sdbscan_amd.h")).read()
    return sorted(set(re.findall(r"\b(sdb_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("sdb_create", "sdb_build_index", "sdb_find_core", "sdb_form_clusters",
              "sdb_core_dist", "sdb_optics_reach", "sdb_get_index"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sdbscan_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    assert set(declared_symbols()) == set(sdbscan_amd._SIGS)


def test_dropin_module_surface():
    m = sdbscan_amd.module()
    cls = m.sDbscan
    for name in ("set_params", "set_sngParams", "set_kernel_params", "clear", "set_topM",
                 "set_topK", "set_threads", "set_clusterNoise", "fit_sDbscan", "test_sDbscan",
                 "fit_sOptics", "load_fit_sDbscan", "load_test_sDbscan", "load_fit_sOptics",
                 "fit_sngDbscan", "test_sngDbscan", "load_test_sngDbscan", "fit_sngOptics", "load_fit_sngOptics",
                 "labels_", "n_clusters_", "ordering_", "reachability_"):
        assert hasattr(cls, name), name
    doc = cls.set_params.__doc__
    for kw, val in (("n_proj", "1024"), ("topK", "10"), ("topM", "50"), ("distance", "'Cosine'"),
                    ("ker_n_features", "1024"), ("ker_sigma", "1.0"), ("clusterNoise", "0"),
                    ("verbose", "False"), ("n_threads", "-1"), ("random_seed", "-1")):
        assert re.search(rf"\b{kw}: [^=,]*= {re.escape(val)}", doc), kw
    obj = cls(100, 8)
    obj.set_params(n_proj=64, topK=3, topM=10, random_seed=1)
    assert obj.labels_ == [] and obj.n_clusters_ == 0


def test_host_rng_matches_oracle_kats():
    """sdb_hd3_bits / sdb_rff_matrix (host std:: calls inside the library)
    give the oracle's bitHD3Generator / gaussGenerator / cauchyGenerator
    draws for the same seeds."""
    import numpy as np
    import sdbscan_amd
    from oracle import oracle as O
    for seed in (0, 42, 12345):
        assert np.array_equal(sdbscan_amd.hd3_bits(3 * 1024, seed), O.hd3_bits(3 * 1024, seed))
    r = sdbscan_amd.rff_matrix(sdbscan_amd.L2, 512, 51, 2.0, 7)
    assert np.array_equal(r, np.asarray(O.gauss(512, 51, 0.0, 0.5, 7)).ravel())
    r = sdbscan_amd.rff_matrix(sdbscan_amd.L1, 64, 13, 1.0, 3)
    assert np.array_equal(r, np.asarray(O.cauchy(64, 13, 0.0, 1.0, 3)).ravel())


def test_struct_layouts_match_ctypes(tmp_path):
    """sdb_params, sdb_stats and sdb_batch_job as the C compiler lays them out
    (size and every field offset) equal the ctypes mirrors."""
    import subprocess
    structs = {"sdb_params": sdbscan_amd.Params, "sdb_stats": sdbscan_amd.Stats,
               "sdb_batch_job": sdbscan_amd.BatchJob}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sdbscan_amd.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_cpp_dropin_class_builds_and_configures(tmp_path):
    """A C++ caller of the drop-in class (INTEGRATION: one changed line per
    fit call) compiles against sDbscan.h, links sDbscan.cpp and the C ABI
    library, and runs the host-side configuration (no device call before a
    fit)."""
    import subprocess
    src = tmp_path / "caller.cpp"
    src.write_text('#include <cstdio>\n#include "sDbscan.h"\n'
                   'int main() { sDbscan dbs(100, 8); dbs.set_params(64, 5, 20, "Cosine");\n'
                   '  dbs.set_precision("BF16"); std::printf("configured\\n"); return 0; }\n')
    exe = tmp_path / "caller"
    lib_dir = os.path.dirname(sdbscan_amd.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "sdbscan_amd", "csrc"), str(src),
                    os.path.join(ROOT, "sdbscan_amd", "csrc", "sDbscan.cpp"), "-L", lib_dir, "-lsdbscan_amd",
                    "-Wl,-rpath," + lib_dir, "-fopenmp", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.strip() == "configured"


def test_product_path_fails_loudly_without_device_or_library():
    """No CPU fallback: without a visible HIP device the context refuses to
    start (SDB_EHIP), and without the built library the package raises."""
    import subprocess
    import sys
    code = (
        "import sdbscan_amd as S\n"
        "try:\n"
        "    S.Context()\n"
        "    print('created')\n"
        "except S.SdbError as e:\n"
        "    print('refused', e.code)\n"
    )
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True,
                         text=True, timeout=120)
    assert out.stdout.strip() == "refused 2", out.stdout + out.stderr
    code = (
        "import sdbscan_amd as S\n"
        "S.LIB_PATH = '/nonexistent/libsdbscan_amd.so'\n"
        "try:\n"
        "    S.capi()\n"
        "    print('loaded')\n"
        "except RuntimeError as e:\n"
        "    print('missing' if 'not built' in str(e) else str(e))\n"
    )
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                         timeout=120)
    assert out.stdout.strip() == "missing", out.stdout + out.stderr


def test_eps_sweep_rounds_like_the_reference(tmp_path):
    """test_sDbscan / test_sngDbscan sweep eps + 1.0 * i * rangeEps in double
    and round once to float (src/sDbscan.cpp:1016, :1221).  A float-only
    expression differs by one ulp for a few inputs in a hundred, e.g.
    0.7f + 3 * 0.05f -> 0.849999964 (double, then float) against 0.850000024
    (float throughout)."""
    import subprocess
    src = tmp_path / "sweep.cpp"
    src.write_text(
        '#include <cstdio>\n#include <random>\n#include "sDbscan.h"\n'
        'int main() {\n'
        '  int bad = 0, float_differs = 0;\n'
        '  std::mt19937 g(7); std::uniform_real_distribution<float> u(0.0f, 1.0f);\n'
        '  for (int t = 0; t < 20000; ++t) {\n'
        '    const float eps = u(g), range = 0.1f * u(g); const int i = t % 16;\n'
        '    const float ref = eps + 1.0 * i * range;  /* the reference statement */\n'
        '    bad += sDbscan::sweep_eps(eps, range, i) != ref;\n'
        '    float_differs += (eps + 1.0f * i * range) != ref;\n'
        '  }\n'
        '  std::printf("%d %d %.9g\\n", bad, float_differs, sDbscan::sweep_eps(0.7f, 0.05f, 3));\n'
        '  return 0; }\n')
    exe = tmp_path / "sweep"
    lib_dir = os.path.dirname(sdbscan_amd.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "sdbscan_amd", "csrc"), str(src), "-L", lib_dir, "-lsdbscan_amd",
                    "-Wl,-rpath," + lib_dir, "-o", str(exe)], check=True)
    bad, differs, e = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(bad) == 0
    assert int(differs) > 100  # the float-only form would not pass
    assert e == "0.849999964"
