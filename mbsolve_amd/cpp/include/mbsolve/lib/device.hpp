/*
 * Device geometry: regions of materials along x
 * (reference: mbsolve-lib/include/mbsolve/lib/device.hpp,
 * mbsolve-lib/src/device.cpp).
 */
#ifndef MBSOLVE_LIB_DEVICE_H
#define MBSOLVE_LIB_DEVICE_H

#include <memory>
#include <set>
#include <string>
#include <vector>
#include <mbsolve/lib/material.hpp>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/** Boundary condition of the field (base class). */
class bc_field
{
public:
    virtual ~bc_field() {}
};

/** Reflectivity boundary condition.  Like in the reference solvers it is
 * stored but not evaluated: the ends are perfect magnetic conductors
 * (H = 0). */
class bc_field_reflectivity : public bc_field
{
private:
    real m_refl_left;
    real m_refl_right;

public:
    explicit bc_field_reflectivity(real refl_left = 1.0,
                                   real refl_right = 1.0)
      : m_refl_left(refl_left), m_refl_right(refl_right)
    {}

    real get_refl_right() { return m_refl_right; }

    real get_refl_left() { return m_refl_left; }
};

/** Section [x_start, x_end] of the device filled with one material. */
class region
{
private:
    std::string m_name;
    std::shared_ptr<material> m_mat;
    real m_x_start;
    real m_x_end;

public:
    region(
        const std::string& name,
        std::shared_ptr<material> mat,
        real x_start,
        real x_end)
      : m_name(name), m_mat(mat), m_x_start(x_start), m_x_end(x_end)
    {}

    const std::string& get_name() const { return m_name; }

    real get_length() const { return m_x_end - m_x_start; }

    real get_x_start() const { return m_x_start; }

    real get_x_end() const { return m_x_end; }

    std::shared_ptr<material> get_material() const { return m_mat; }
};

/** A device: ordered regions and the set of material ids they use. */
class device
{
private:
    std::string m_name;
    std::vector<std::shared_ptr<region> > m_regions;
    std::set<std::string> m_used_materials;
    std::shared_ptr<bc_field> m_bc_field;

public:
    device(
        const std::string& name,
        std::shared_ptr<bc_field> field_boundary =
            std::make_shared<bc_field_reflectivity>(1.0, 1.0));

    ~device();

    void add_region(std::shared_ptr<region> reg);

    const std::vector<std::shared_ptr<region> >& get_regions() const;

    /** Material ids in lexicographic order (std::set); this order defines
     * the solvers' material indices. */
    const std::set<std::string>& get_used_materials() const;

    const std::string& get_name() const;

    /** Sum of the region lengths. */
    real get_length() const;

    /** Smallest relative permittivity of all regions (1e42 if none). */
    real get_minimum_permittivity() const;

    std::shared_ptr<bc_field> get_bc_field() const;

    void set_bc_field(std::shared_ptr<bc_field> boundary_field);
};
}

#endif
