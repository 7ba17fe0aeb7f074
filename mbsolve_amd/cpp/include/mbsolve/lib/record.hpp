/*
 * Record requests (reference: mbsolve-lib/include/mbsolve/lib/record.hpp,
 * mbsolve-lib/src/record.cpp).
 */
#ifndef MBSOLVE_LIB_RECORD_H
#define MBSOLVE_LIB_RECORD_H

#include <string>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/** Which trace to store, at which interval and position. */
class record
{
public:
    enum type
    {
        electric,
        polar_dt,
        magnetic,
        density,
        inversion
    };

private:
    std::string m_name;
    type m_type;
    unsigned int m_col;
    unsigned int m_row;
    real m_position;
    real m_interval;
    bool m_is_complex;

public:
    /** Type deduced from the first letter of the name: e, p, h, d, i.
     * interval 0: every step; position < 0: the whole grid.  Unlike the
     * reference (record.cpp:27-56, which leaves them uninitialised), the
     * row/column indices are set to 0. */
    record(const std::string& name, real interval = 0.0,
           real position = -1.0);

    /** Explicit type; row_idx and col_idx are one-based. */
    record(
        const std::string& name,
        type record_type,
        unsigned int row_idx,
        unsigned int col_idx,
        real interval = 0.0,
        real position = -1.0);

    const std::string& get_name() const { return m_name; }

    type get_type() const { return m_type; }

    unsigned int get_col() const { return m_col; }

    unsigned int get_row() const { return m_row; }

    real get_position() const { return m_position; }

    real get_interval() const { return m_interval; }

    bool is_complex() const { return m_is_complex; }
};
}

#endif
