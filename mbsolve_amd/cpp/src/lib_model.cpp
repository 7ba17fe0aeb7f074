/*
 * Implementation of the mbsolve physical-model API: relaxation
 * superoperator, two-level description, materials, devices, records,
 * results and scenarios.
 */
#include <algorithm>
#include <cmath>
#include <iterator>
#include <stdexcept>
#include <mbsolve/lib.hpp>

namespace mbsolve {

/* ---- qm_lindblad_relaxation (qm_description.hpp:166-236) ---------------- */
qm_lindblad_relaxation::qm_lindblad_relaxation(
    const std::vector<std::vector<real> >& scattering_rates,
    const std::vector<real>& pure_deph_rates)
  : qm_superop(scattering_rates.size()), m_scattering(scattering_rates),
    m_dephasing(pure_deph_rates)
{
    const unsigned int n = m_num_levels;
    for (unsigned int m = 0; m < n; m++) {
        if (m_scattering[m].size() != n) {
            throw std::invalid_argument(
                "Scattering rate matrix must be square.");
        }
    }
    if (m_dephasing.size() < n * (n - 1) / 2) {
        throw std::invalid_argument(
            "Too few pure dephasing rates for the number of levels.");
    }
    /* main diagonal: minus the inverse population life time of level m,
     * i.e. the sum of all rates out of m */
    for (unsigned int m = 0; m < n; m++) {
        real out = 0;
        for (unsigned int j = 0; j < n; j++) {
            if (j != m) {
                out += m_scattering[j][m];
            }
        }
        m_scattering[m][m] = -out;
    }
    /* total dephasing of coherence (m, n), column-major upper triangle */
    unsigned int idx = 0;
    for (unsigned int col = 0; col < n; col++) {
        for (unsigned int row = 0; row < col; row++) {
            real g = -std::abs(m_dephasing[idx]);
            g -= 0.5 *
                (std::abs(m_scattering[row][row]) +
                 std::abs(m_scattering[col][col]));
            m_dephasing[idx] = g;
            idx++;
        }
    }
}

qm_operator
qm_lindblad_relaxation::action(const qm_operator& arg) const
{
    const unsigned int n = m_num_levels;
    std::vector<real> pop(n);
    for (unsigned int m = 0; m < n; m++) {
        real acc = 0;
        for (unsigned int j = 0; j < n; j++) {
            acc += m_scattering[m][j] * arg.get_main_diagonal()[j];
        }
        pop[m] = acc;
    }
    std::vector<std::complex<real> > coh = arg.get_off_diagonal();
    for (size_t i = 0; i < coh.size(); i++) {
        coh[i] *= m_dephasing[i];
    }
    return qm_operator(pop, coh);
}

/* ---- qm_desc_2lvl (qm_description.hpp:375-404) -------------------------- */
qm_desc_2lvl::qm_desc_2lvl(
    real carrier_density,
    real transition_freq,
    real dipole_moment,
    real scattering_rate,
    real dephasing_rate,
    real equilibrium_inversion)
  : qm_description(
        carrier_density,
        qm_operator(
            { -HBAR * transition_freq / 2, +HBAR * transition_freq / 2 }),
        qm_operator({ 0, 0 }, { -E0 * dipole_moment }),
        std::make_shared<qm_lindblad_relaxation>(
            std::vector<std::vector<real> >{
                { 0, scattering_rate * 0.5 * (1 - equilibrium_inversion) },
                { scattering_rate * 0.5 * (1 + equilibrium_inversion), 0 } },
            std::vector<real>{ dephasing_rate - 0.5 * scattering_rate })),
    m_trans_freq(transition_freq), m_dipole_mom(dipole_moment),
    m_scattering(scattering_rate), m_dephasing(dephasing_rate),
    m_equi_inv(equilibrium_inversion)
{}

/* ---- material library (material.cpp:29-47) ------------------------------ */
std::map<std::string, std::shared_ptr<material> >&
material::library()
{
    static std::map<std::string, std::shared_ptr<material> > lib;
    return lib;
}

void
material::add_to_library(std::shared_ptr<material> mat)
{
    if (!mat) {
        throw std::invalid_argument("Material is null.");
    }
    auto& lib = library();
    if (lib.find(mat->get_id()) != lib.end()) {
        throw std::invalid_argument("Material already available in library.");
    }
    lib[mat->get_id()] = mat;
}

std::shared_ptr<material>
material::get_from_library(const std::string& id)
{
    auto& lib = library();
    auto it = lib.find(id);
    if (it == lib.end()) {
        throw std::invalid_argument("Material not available in library.");
    }
    return it->second;
}

void
material::clear_library()
{
    library().clear();
}

/* ---- device (device.cpp) ------------------------------------------------ */
device::device(const std::string& name,
               std::shared_ptr<bc_field> field_boundary)
  : m_name(name), m_bc_field(field_boundary)
{}

device::~device() {}

void
device::add_region(std::shared_ptr<region> reg)
{
    if (!reg || !reg->get_material()) {
        throw std::invalid_argument("Region or its material is null.");
    }
    m_regions.push_back(reg);
    m_used_materials.insert(reg->get_material()->get_id());
}

const std::vector<std::shared_ptr<region> >&
device::get_regions() const
{
    return m_regions;
}

const std::set<std::string>&
device::get_used_materials() const
{
    return m_used_materials;
}

const std::string&
device::get_name() const
{
    return m_name;
}

real
device::get_length() const
{
    real total = 0.0;
    for (const auto& r : m_regions) {
        total += r->get_length();
    }
    return total;
}

real
device::get_minimum_permittivity() const
{
    real lowest = 1e42;
    for (const auto& r : m_regions) {
        lowest = std::min(lowest, r->get_material()->get_rel_permittivity());
    }
    return lowest;
}

std::shared_ptr<bc_field>
device::get_bc_field() const
{
    return m_bc_field;
}

void
device::set_bc_field(std::shared_ptr<bc_field> boundary_field)
{
    m_bc_field = boundary_field;
}

/* ---- record (record.cpp:27-69) ------------------------------------------ */
record::record(const std::string& name, real interval, real position)
  : m_name(name), m_col(0), m_row(0), m_position(position),
    m_interval(interval)
{
    if (name.empty()) {
        throw std::invalid_argument("Unknown record type");
    }
    switch (name[0]) {
    case 'e':
        m_type = type::electric;
        break;
    case 'p':
        m_type = type::polar_dt;
        break;
    case 'h':
        m_type = type::magnetic;
        break;
    case 'd':
        m_type = type::density;
        break;
    case 'i':
        m_type = type::inversion;
        break;
    default:
        throw std::invalid_argument("Unknown record type");
    }
    m_is_complex = (m_type == type::density) && (m_row != m_col);
}

record::record(
    const std::string& name,
    type record_type,
    unsigned int row_idx,
    unsigned int col_idx,
    real interval,
    real position)
  : m_name(name), m_type(record_type), m_col(col_idx - 1), m_row(row_idx - 1),
    m_position(position), m_interval(interval)
{
    m_is_complex = (m_type == type::density) && (m_row != m_col);
}

/* ---- result ------------------------------------------------------------- */
std::vector<std::complex<real> >
result::get_data_complex() const
{
    std::vector<std::complex<real> > out;
    out.reserve(m_real.size());
    for (size_t i = 0; i < m_real.size(); i++) {
        out.emplace_back(m_real[i], m_imag[i]);
    }
    return out;
}

/* ---- scenario (scenario.cpp) -------------------------------------------- */
scenario::scenario(
    const std::string& name,
    unsigned int num_gridpoints,
    real endtime,
    const qm_operator& rho_init,
    unsigned int num_timesteps)
  : scenario(name, num_gridpoints, endtime,
             std::make_shared<ic_density_const>(rho_init),
             std::make_shared<ic_field_random>(),
             std::make_shared<ic_field_const>(0), num_timesteps)
{}

scenario::scenario(
    const std::string& name,
    unsigned int num_gridpoints,
    real endtime,
    std::shared_ptr<ic_density> density_init,
    std::shared_ptr<ic_field> electric_init,
    std::shared_ptr<ic_field> magnetic_init,
    unsigned int num_timesteps)
  : m_name(name), m_num_timesteps(num_timesteps),
    m_num_gridpoints(num_gridpoints), m_timestep_size(0.0),
    m_gridpoint_size(0.0), m_endtime(endtime), m_dens_init(density_init),
    m_h_init(magnetic_init), m_e_init(electric_init)
{}

void scenario::add_record(std::shared_ptr<record> rec)
{
    m_records.push_back(rec);
}
const std::vector<std::shared_ptr<record> >& scenario::get_records() const
{
    return m_records;
}
void scenario::add_source(std::shared_ptr<source> src)
{
    m_sources.push_back(src);
}
const std::vector<std::shared_ptr<source> >& scenario::get_sources() const
{
    return m_sources;
}
const std::string& scenario::get_name() const { return m_name; }
unsigned int scenario::get_num_timesteps() const { return m_num_timesteps; }
void scenario::set_num_timesteps(unsigned int v) { m_num_timesteps = v; }
unsigned int scenario::get_num_gridpoints() const { return m_num_gridpoints; }
void scenario::set_num_gridpoints(unsigned int v) { m_num_gridpoints = v; }
real scenario::get_timestep_size() const { return m_timestep_size; }
void scenario::set_timestep_size(real v) { m_timestep_size = v; }
real scenario::get_gridpoint_size() const { return m_gridpoint_size; }
void scenario::set_gridpoint_size(real v) { m_gridpoint_size = v; }
real scenario::get_endtime() const { return m_endtime; }
void scenario::set_endtime(real v) { m_endtime = v; }
std::shared_ptr<ic_density> scenario::get_ic_density() const
{
    return m_dens_init;
}
void scenario::set_ic_density(std::shared_ptr<ic_density> v)
{
    m_dens_init = v;
}
std::shared_ptr<ic_field> scenario::get_ic_electric() const
{
    return m_e_init;
}
void scenario::set_ic_electric(std::shared_ptr<ic_field> v) { m_e_init = v; }
std::shared_ptr<ic_field> scenario::get_ic_magnetic() const
{
    return m_h_init;
}
void scenario::set_ic_magnetic(std::shared_ptr<ic_field> v) { m_h_init = v; }

/* ---- solver factory (solver.cpp) ---------------------------------------- */
std::map<std::string, solver::bootstrap_t>&
solver::bootstraps()
{
    static std::map<std::string, bootstrap_t> b;
    return b;
}

std::shared_ptr<solver>
solver::create_instance(
    const std::string& name,
    std::shared_ptr<const device> dev,
    std::shared_ptr<scenario> scen)
{
    auto it = bootstraps().find(name);
    if (it == bootstraps().end()) {
        throw std::invalid_argument("Unknown solver " + name);
    }
    return it->second(dev, scen);
}

void
solver::register_bootstrap(const std::string& name, bootstrap_t b)
{
    auto& m = bootstraps();
    if (m.find(name) != m.end()) {
        throw std::invalid_argument("Solver already registered.");
    }
    m[name] = b;
}

solver::solver(
    const std::string& name,
    std::shared_ptr<const device> dev,
    std::shared_ptr<scenario> scen)
  : m_name(name), m_device(dev), m_scenario(scen)
{}

solver::~solver() {}

const std::string& solver::get_name() const { return m_name; }
const scenario& solver::get_scenario() const { return *m_scenario; }
const device& solver::get_device() const { return *m_device; }
const std::vector<std::shared_ptr<result> >& solver::get_results() const
{
    return m_results;
}

std::vector<std::string>
solver::get_avail_solvers()
{
    std::vector<std::string> names;
    for (const auto& kv : bootstraps()) {
        names.push_back(kv.first);
    }
    return names;
}

/* ---- writer factory (writer.cpp) ---------------------------------------- */
std::map<std::string, writer::bootstrap_t>&
writer::bootstraps()
{
    static std::map<std::string, bootstrap_t> b;
    return b;
}

std::shared_ptr<writer>
writer::create_instance(const std::string& name)
{
    auto it = bootstraps().find(name);
    if (it == bootstraps().end()) {
        throw std::invalid_argument("Unknown writer " + name);
    }
    return it->second();
}

void
writer::register_bootstrap(const std::string& name, bootstrap_t b)
{
    auto& m = bootstraps();
    if (m.find(name) != m.end()) {
        throw std::invalid_argument("Writer already registered.");
    }
    m[name] = b;
}

writer::writer(const std::string& name, const std::string& extension)
  : m_name(name), m_ext(extension)
{}

writer::~writer() {}

const std::string& writer::get_name() const { return m_name; }
const std::string& writer::get_extension() const { return m_ext; }

std::vector<std::string>
writer::get_avail_writers()
{
    std::vector<std::string> names;
    for (const auto& kv : bootstraps()) {
        names.push_back(kv.first);
    }
    return names;
}
}
