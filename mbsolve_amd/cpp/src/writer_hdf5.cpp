/*
 * HDF5 result writer on the HDF5 C API (reference:
 * writer-hdf5/src/writer_hdf5.cpp:31-109 for the file layout).
 */
#include <mutex>
#include <stdexcept>
#include <hdf5.h>
#include <mbsolve/writer-hdf5/writer_hdf5.hpp>

namespace mbsolve {

namespace {

/* owning handle: closes with the matching H5*close on scope exit */
class h5id
{
    hid_t m_id;
    herr_t (*m_close)(hid_t);

public:
    h5id(hid_t id, herr_t (*close)(hid_t), const char* what)
      : m_id(id), m_close(close)
    {
        if (id < 0) {
            throw std::runtime_error(std::string("HDF5: ") + what +
                                     " failed");
        }
    }
    ~h5id() { m_close(m_id); }
    h5id(const h5id&) = delete;
    h5id& operator=(const h5id&) = delete;
    operator hid_t() const { return m_id; }
};

void
check(herr_t e, const char* what)
{
    if (e < 0) {
        throw std::runtime_error(std::string("HDF5: ") + what + " failed");
    }
}

void
scalar_attribute(hid_t loc, const char* name, hid_t type, const void* v)
{
    h5id space(H5Screate(H5S_SCALAR), H5Sclose, "H5Screate");
    h5id attr(H5Acreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT),
              H5Aclose, name);
    check(H5Awrite(attr, type, v), name);
}

void
dataset_2d(hid_t group, const char* name, unsigned long long rows,
           unsigned long long cols, const double* data)
{
    const hsize_t dims[2] = { rows, cols };
    h5id space(H5Screate_simple(2, dims, nullptr), H5Sclose,
               "H5Screate_simple");
    h5id ds(H5Dcreate2(group, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT,
                       H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose, name);
    if (rows * cols) {
        check(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                       data),
              name);
    }
}

std::vector<double>
read_dataset(hid_t group, const char* name, unsigned long long& rows,
             unsigned long long& cols)
{
    h5id ds(H5Dopen2(group, name, H5P_DEFAULT), H5Dclose, name);
    h5id space(H5Dget_space(ds), H5Sclose, "H5Dget_space");
    hsize_t dims[2] = { 0, 0 };
    if (H5Sget_simple_extent_ndims(space) != 2) {
        throw std::runtime_error(std::string("HDF5: dataset ") + name +
                                 " is not two-dimensional");
    }
    H5Sget_simple_extent_dims(space, dims, nullptr);
    rows = dims[0];
    cols = dims[1];
    std::vector<double> v(rows * cols);
    if (!v.empty()) {
        check(H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                      v.data()),
              name);
    }
    return v;
}

herr_t
collect_name(hid_t, const char* name, const H5L_info_t*, void* op)
{
    static_cast<std::vector<std::string>*>(op)->push_back(name);
    return 0;
}

/* the HDF5 library is used from one thread at a time; its own error
 * printing is off (failures surface as exceptions) */
std::mutex g_h5_mutex;

void
quiet()
{
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}
}

writer_hdf5::writer_hdf5() : writer("hdf5", "hdf") {}

writer_hdf5::~writer_hdf5() {}

void
writer_hdf5::write(
    const std::string& filename,
    const std::vector<std::shared_ptr<result> >& results,
    std::shared_ptr<const device> dev,
    std::shared_ptr<const scenario> scen) const
{
    std::lock_guard<std::mutex> lock(g_h5_mutex);
    quiet();
    h5id file(H5Fcreate(filename.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT,
                        H5P_DEFAULT),
              H5Fclose, "H5Fcreate");
    const double timestep = scen->get_timestep_size();
    const double gridpoint = scen->get_gridpoint_size();
    const double endtime = scen->get_endtime();
    const double length = dev->get_length();
    scalar_attribute(file, "timestep_size", H5T_NATIVE_DOUBLE, &timestep);
    scalar_attribute(file, "gridpoint_size", H5T_NATIVE_DOUBLE, &gridpoint);
    scalar_attribute(file, "sim_endtime", H5T_NATIVE_DOUBLE, &endtime);
    scalar_attribute(file, "dev_length", H5T_NATIVE_DOUBLE, &length);
    for (const auto& res : results) {
        h5id group(H5Gcreate2(file, res->get_name().c_str(), H5P_DEFAULT,
                              H5P_DEFAULT, H5P_DEFAULT),
                   H5Gclose, "H5Gcreate2");
        const hbool_t cplx = res->is_complex();
        scalar_attribute(group, "is_complex", H5T_NATIVE_HBOOL, &cplx);
        dataset_2d(group, "real", res->get_rows(), res->get_cols(),
                   res->get_data_real().data());
        if (cplx) {
            dataset_2d(group, "imag", res->get_rows(), res->get_cols(),
                       res->get_data_imag().data());
        }
    }
}

writer_hdf5_loader::writer_hdf5_loader()
{
    static std::once_flag once;
    std::call_once(once,
                   []() { writer::register_writer<writer_hdf5>("hdf5"); });
}

hdf5_file_data
read_hdf5(const std::string& filename)
{
    std::lock_guard<std::mutex> lock(g_h5_mutex);
    quiet();
    hdf5_file_data out;
    h5id file(H5Fopen(filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT),
              H5Fclose, "H5Fopen");
    for (const char* a : { "timestep_size", "gridpoint_size", "sim_endtime",
                           "dev_length" }) {
        if (H5Aexists(file, a) > 0) {
            h5id attr(H5Aopen(file, a, H5P_DEFAULT), H5Aclose, a);
            double v = 0.0;
            check(H5Aread(attr, H5T_NATIVE_DOUBLE, &v), a);
            out.attributes[a] = v;
        }
    }
    std::vector<std::string> names;
    check(H5Literate(file, H5_INDEX_NAME, H5_ITER_INC, nullptr, collect_name,
                     &names),
          "H5Literate");
    for (const auto& nm : names) {
        h5id group(H5Gopen2(file, nm.c_str(), H5P_DEFAULT), H5Gclose,
                   "H5Gopen2");
        hdf5_result_data r;
        h5id attr(H5Aopen(group, "is_complex", H5P_DEFAULT), H5Aclose,
                  "is_complex");
        hbool_t c = false;
        check(H5Aread(attr, H5T_NATIVE_HBOOL, &c), "is_complex");
        r.is_complex = c;
        r.real = read_dataset(group, "real", r.rows, r.cols);
        if (r.is_complex) {
            unsigned long long ri = 0, ci = 0;
            r.imag = read_dataset(group, "imag", ri, ci);
        }
        out.results[nm] = std::move(r);
    }
    return out;
}
}
