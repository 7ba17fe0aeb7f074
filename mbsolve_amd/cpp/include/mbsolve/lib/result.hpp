/*
 * Result traces (reference: mbsolve-lib/include/mbsolve/lib/result.hpp).
 */
#ifndef MBSOLVE_LIB_RESULT_H
#define MBSOLVE_LIB_RESULT_H

#include <string>
#include <vector>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/** rows (time samples) x cols (grid points), row-major, real and
 * imaginary part. */
class result
{
private:
    std::string m_name;
    unsigned int m_cols;
    unsigned int m_rows;
    unsigned int m_count;
    std::vector<real> m_real;
    std::vector<real> m_imag;
    bool m_is_complex;

public:
    explicit result(
        const std::string& name,
        unsigned int cols,
        unsigned int rows,
        bool is_complex = false)
      : m_name(name), m_cols(cols), m_rows(rows), m_count(cols * rows),
        m_real(cols * rows), m_imag(cols * rows), m_is_complex(is_complex)
    {}

    std::string get_name() const { return m_name; }

    unsigned int get_count() const { return m_count; }

    unsigned int get_cols() const { return m_cols; }

    unsigned int get_rows() const { return m_rows; }

    bool is_complex() const { return m_is_complex; }

    std::vector<real>::iterator get_data_real_it(unsigned int row,
                                                 unsigned int col)
    {
        return m_real.begin() + (size_t)row * m_cols + col;
    }

    std::vector<real>::iterator get_data_imag_it(unsigned int row,
                                                 unsigned int col)
    {
        return m_imag.begin() + (size_t)row * m_cols + col;
    }

    const std::vector<real>& get_data_real() const { return m_real; }

    const std::vector<real>& get_data_imag() const { return m_imag; }

    std::vector<std::complex<real> > get_data_complex() const;
};
}

#endif
