/*
 * HDF5 result writer (reference: writer-hdf5/include/mbsolve/writer-hdf5/
 * writer_hdf5.hpp, writer-hdf5/src/writer_hdf5.cpp).  File layout, as the
 * reference writes it:
 *   root attributes (FP64 scalars) timestep_size, gridpoint_size,
 *   sim_endtime, dev_length;
 *   one group per result, named after it, with the scalar attribute
 *   is_complex (hbool) and the dataset "real" (FP64, [rows][cols]), plus
 *   "imag" of the same shape when the result is complex.
 */
#ifndef MBSOLVE_WRITER_HDF5_H
#define MBSOLVE_WRITER_HDF5_H

#include <map>
#include <string>
#include <vector>
#include <mbsolve/lib/writer.hpp>

namespace mbsolve {

class writer_hdf5 : public writer
{
public:
    writer_hdf5();

    ~writer_hdf5();

    void write(
        const std::string& file,
        const std::vector<std::shared_ptr<result> >& results,
        std::shared_ptr<const device> dev,
        std::shared_ptr<const scenario> scen) const override;
};

/** Registers the writer "hdf5" (writer_hdf5_loader of the reference). */
class writer_hdf5_loader
{
public:
    writer_hdf5_loader();
};

/** Contents of a file in the layout above (for tests and post-processing
 * without an HDF5 binding). */
struct hdf5_result_data
{
    bool is_complex = false;
    unsigned long long rows = 0, cols = 0;
    std::vector<double> real, imag;
};
struct hdf5_file_data
{
    std::map<std::string, double> attributes;
    std::map<std::string, hdf5_result_data> results;
};

hdf5_file_data read_hdf5(const std::string& file);
}

#endif
