/*
 * Python module of the HDF5 writer (the reference's SWIG module
 * writer-hdf5/swig/writerhdf5.i): load_writer() registers "hdf5" with the
 * writer factory; read_file() returns a file's contents.
 */
#include <cstring>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <mbsolve/writer-hdf5/writer_hdf5.hpp>

namespace py = pybind11;
using namespace mbsolve;

PYBIND11_MODULE(_writerhdf5, m)
{
    m.def("load_writer", []() { writer_hdf5_loader load; });
    m.def("read_file", [](const std::string& name) {
        const hdf5_file_data f = read_hdf5(name);
        py::dict attrs, results;
        for (const auto& a : f.attributes) {
            attrs[py::str(a.first)] = a.second;
        }
        for (const auto& r : f.results) {
            auto arr = [&r](const std::vector<double>& v) {
                py::array_t<double> a({ (size_t)r.second.rows,
                                        (size_t)r.second.cols });
                if (!v.empty()) {
                    std::memcpy(a.mutable_data(), v.data(),
                                v.size() * sizeof(double));
                }
                return a;
            };
            py::dict d;
            d["is_complex"] = r.second.is_complex;
            d["real"] = arr(r.second.real);
            if (r.second.is_complex) {
                d["imag"] = arr(r.second.imag);
            }
            results[py::str(r.first)] = d;
        }
        py::dict out;
        out["attributes"] = attrs;
        out["results"] = results;
        return out;
    });
}
