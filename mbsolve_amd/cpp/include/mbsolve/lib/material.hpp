/*
 * Materials and the global material library
 * (reference: mbsolve-lib/include/mbsolve/lib/material.hpp,
 * mbsolve-lib/src/material.cpp).
 */
#ifndef MBSOLVE_LIB_MATERIAL_H
#define MBSOLVE_LIB_MATERIAL_H

#include <map>
#include <memory>
#include <string>
#include <mbsolve/lib/qm_description.hpp>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/** Electromagnetic properties plus an optional quantum-mechanical
 * description; defaults to vacuum. */
class material
{
private:
    std::string m_id;
    std::shared_ptr<qm_description> m_qm;
    real m_rel_permittivity;
    real m_rel_permeability;
    real m_losses;
    real m_overlap_factor;

    static std::map<std::string, std::shared_ptr<material> >& library();

public:
    explicit material(
        const std::string& id,
        std::shared_ptr<qm_description> qm = nullptr,
        real rel_permittivity = 1.0,
        real overlap_factor = 1.0,
        real losses = 0.0,
        real rel_permeability = 1.0)
      : m_id(id), m_qm(qm), m_rel_permittivity(rel_permittivity),
        m_rel_permeability(rel_permeability), m_losses(losses),
        m_overlap_factor(overlap_factor)
    {}

    const std::string& get_id() const { return m_id; }

    std::shared_ptr<qm_description> get_qm() const { return m_qm; }

    real get_rel_permittivity() const { return m_rel_permittivity; }

    real get_rel_permeability() const { return m_rel_permeability; }

    real get_losses() const { return m_losses; }

    real get_overlap_factor() const { return m_overlap_factor; }

    /** Adds a material; throws std::invalid_argument if the id exists. */
    static void add_to_library(std::shared_ptr<material> mat);

    /** Looks a material up; throws std::invalid_argument if unknown. */
    static std::shared_ptr<material> get_from_library(const std::string& id);

    /** Removes every material (not in the reference; lets one process set
     * up several independent scenarios with reused ids). */
    static void clear_library();
};
}

#endif
