// pybind11 module `sDbscan`: drop-in for the reference module
// (python/python_wrapper.cpp:14-65) with the same class, methods, keyword
// names and defaults.  X is accepted as an n_features x n_points array and
// converted to Fortran order float32, which is what the reference's
// Eigen::Ref<const MatrixXf> conversion does.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "sDbscan.h"

namespace py = pybind11;

namespace {

typedef py::array_t<float, py::array::f_style | py::array::forcecast> FArr;

const float* check_x(const sDbscan& self, const FArr& X, int d, int n) {
    if (X.ndim() != 2 || X.shape(0) != d || X.shape(1) != n)
        throw std::invalid_argument("X must be an n_features x n_points matrix (" + std::to_string(d) + " x " +
                                    std::to_string(n) + ")");
    (void)self;
    return X.data();
}

// The class does not expose n_points/n_features; keep them beside it.
struct Dims {
    int n, d;
};

}  // namespace

class PySDbscan : public sDbscan {
public:
    Dims dims;
    PySDbscan(int n, int d) : sDbscan(n, d), dims{n, d} {}
};

PYBIND11_MODULE(sDbscan, m) {
    m.doc() = "sDbscan neighbourhood discovery on AMD Instinct MI355X (HIP, gfx950)";
    py::class_<PySDbscan>(m, "sDbscan")
        .def(py::init<const int&, const int&>(), py::arg("n_points"), py::arg("n_features"))
        .def("set_params", &PySDbscan::set_params, py::arg("n_proj") = 1024, py::arg("topK") = 10,
             py::arg("topM") = 50, py::arg("distance") = "Cosine", py::arg("ker_n_features") = 1024,
             py::arg("ker_sigma") = 1.0, py::arg("ker_intervalSampling") = 0.4,
             py::arg("clusterNoise") = 0, py::arg("samplingProb") = 0.01, py::arg("verbose") = false,
             py::arg("output") = "", py::arg("n_threads") = -1, py::arg("random_seed") = -1)
        .def("set_sngParams", &PySDbscan::set_sngParams, py::arg("distance") = "Cosine",
             py::arg("samplingProb") = 0.01, py::arg("clusterNoise") = 0, py::arg("verbose") = false,
             py::arg("output") = "", py::arg("n_threads") = -1, py::arg("random_seed") = -1)
        .def_readonly("labels_", &PySDbscan::labels_)
        .def_readonly("n_clusters_", &PySDbscan::n_clusters_)
        .def_readonly("ordering_", &PySDbscan::ordering_)
        .def_readonly("reachability_", &PySDbscan::reachability_)
        .def("clear", &PySDbscan::clear)
        .def("set_topM", &PySDbscan::set_topM, py::arg("topM"))
        .def("set_threads", &PySDbscan::set_threads, py::arg("n_threads"))
        .def("set_clusterNoise", &PySDbscan::set_clusterNoise, py::arg("clusterNoise"))
        .def("set_topK", &PySDbscan::set_topK, py::arg("topK"))
        .def("set_kernel_params", &PySDbscan::set_kernel_params, py::arg("distance") = "L2",
             py::arg("ker_n_features") = 1024, py::arg("ker_sigma") = 1.0,
             py::arg("ker_intervalSampling") = 0.4)
        .def(
            "fit_sDbscan",
            [](PySDbscan& s, const FArr& X, float eps, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.fit_sDbscan(p, eps, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("minPts"))
        .def(
            "test_sDbscan",
            [](PySDbscan& s, const FArr& X, float eps, float rangeEps, int n_tests, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.test_sDbscan(p, eps, rangeEps, n_tests, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("rangeEps"), py::arg("n_tests"), py::arg("minPts"))
        .def(
            "fit_sOptics",
            [](PySDbscan& s, const FArr& X, float eps, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.fit_sOptics(p, eps, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("minPts"))
        .def(
            "load_fit_sDbscan",
            [](PySDbscan& s, const std::string& dataset, float eps, int minPts) {
                py::gil_scoped_release rel;
                s.load_fit_sDbscan(dataset, eps, minPts);
            },
            py::arg("dataset"), py::arg("eps"), py::arg("minPts"))
        .def(
            "load_test_sDbscan",
            [](PySDbscan& s, const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts) {
                py::gil_scoped_release rel;
                s.load_test_sDbscan(dataset, eps, rangeEps, n_tests, minPts);
            },
            py::arg("dataset"), py::arg("eps"), py::arg("rangeEps"), py::arg("n_tests"), py::arg("minPts"))
        .def(
            "load_fit_sOptics",
            [](PySDbscan& s, const std::string& dataset, float eps, int minPts) {
                py::gil_scoped_release rel;
                s.load_fit_sOptics(dataset, eps, minPts);
            },
            py::arg("dataset"), py::arg("eps"), py::arg("minPts"))
        .def(
            "fit_sngDbscan",
            [](PySDbscan& s, const FArr& X, float eps, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.fit_sngDbscan(p, eps, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("minPts"))
        .def(
            "test_sngDbscan",
            [](PySDbscan& s, const FArr& X, float eps, float rangeEps, int n_tests, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.test_sngDbscan(p, eps, rangeEps, n_tests, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("rangeEps"), py::arg("n_tests"), py::arg("minPts"))
        .def(
            "load_test_sngDbscan",
            [](PySDbscan& s, const std::string& dataset, float eps, float rangeEps, int n_tests, int minPts) {
                py::gil_scoped_release rel;
                s.load_test_sngDbscan(dataset, eps, rangeEps, n_tests, minPts);
            },
            py::arg("dataset"), py::arg("eps"), py::arg("rangeEps"), py::arg("n_tests"), py::arg("minPts"))
        .def(
            "fit_sngOptics",
            [](PySDbscan& s, const FArr& X, float eps, int minPts) {
                const float* p = check_x(s, X, s.dims.d, s.dims.n);
                py::gil_scoped_release rel;
                s.fit_sngOptics(p, eps, minPts);
            },
            py::arg("X"), py::arg("eps"), py::arg("minPts"))
        .def(
            "load_fit_sngOptics",
            [](PySDbscan& s, const std::string& dataset, float eps, int minPts) {
                py::gil_scoped_release rel;
                s.load_fit_sngOptics(dataset, eps, minPts);
            },
            py::arg("dataset"), py::arg("eps"), py::arg("minPts"))
        // extensions of this build
        .def("set_precision", &PySDbscan::set_precision, py::arg("precision"))
        .def("set_device", &PySDbscan::set_device, py::arg("device"))
        .def("get_index",
             [](PySDbscan& s) {
                 std::vector<int> tk, tm;
                 s.get_index(tk, tm);
                 const int K = s.get_topK(), M = s.get_topM(), P = s.get_n_proj();
                 py::array_t<int> a({(py::ssize_t)s.dims.n, (py::ssize_t)(2 * K)});
                 std::copy(tk.begin(), tk.end(), a.mutable_data());
                 py::array_t<int> b({(py::ssize_t)P, (py::ssize_t)(2 * M)});
                 std::copy(tm.begin(), tm.end(), b.mutable_data());
                 return py::make_tuple(a, b);
             })
        .def("get_random",
             [](PySDbscan& s) {
                 const auto& bits = s.get_bits();
                 py::array_t<uint8_t> a((py::ssize_t)bits.size());
                 std::copy(bits.begin(), bits.end(), a.mutable_data());
                 const auto& R = s.get_R();
                 py::array_t<float> r((py::ssize_t)R.size());
                 std::copy(R.begin(), R.end(), r.mutable_data());
                 return py::make_tuple(a, r, s.get_fhtDim());
             })
        .def("stats", [](const PySDbscan& s) {
            sdb_stats t = s.get_stats();
            py::dict d;
            d["t_transform_ms"] = t.t_transform_ms;
            d["t_index_ms"] = t.t_index_ms;
            d["t_core_ms"] = t.t_core_ms;
            d["t_cluster_ms"] = t.t_cluster_ms;
            d["t_dist_ms"] = t.t_dist_ms;
            d["n_pairs"] = t.n_pairs;
            d["n_queue"] = t.n_queue;
            d["n_recheck32"] = t.n_recheck32;
            d["n_recheck64"] = t.n_recheck64;
            d["n_recheck_hits"] = t.n_recheck_hits;
            d["n_hubs"] = t.n_hubs;
            d["n_tie_points"] = t.n_tie_points;
            d["n_tie_rows"] = t.n_tie_rows;
            d["n_core"] = t.n_core;
            d["n_components"] = t.n_components;
            d["band"] = t.band;
            d["band32"] = t.band32;
            d["n_near_eps"] = t.n_near_eps;
            return d;
        });
}
