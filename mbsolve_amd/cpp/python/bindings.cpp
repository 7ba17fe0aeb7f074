/*
 * Python bindings of the mbsolve API and the HIP solvers, mirroring the
 * reference's SWIG modules mbsolve.lib / mbsolve.solvercpu
 * (mbsolve-lib/swig/lib.i, solver-cpu/swig/solvercpu.i): C++ exceptions
 * surface as Python exceptions, results as numpy arrays.
 */
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <mbsolve/lib.hpp>
#include <mbsolve/lib/internal/cv_representation.hpp>
#include <mbsolve/lib/internal/fdtd_setup.hpp>
#include <mbsolve/lib/internal/numerics.hpp>
#include <mbsolve/solver-hip/solver_hip_fdtd.hpp>

namespace py = pybind11;
using namespace mbsolve;

namespace {

template<typename T>
py::array_t<T> to_np(const std::vector<T>& v)
{
    py::array_t<T> a(v.size());
    if (!v.empty()) {
        std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
    }
    return a;
}

py::array_t<double> mat_to_np(const num::rmat& m)
{
    py::array_t<double> a({ m.rows, m.cols });
    std::memcpy(a.mutable_data(), m.a.data(), m.a.size() * sizeof(double));
    return a;
}

num::rmat np_to_mat(py::array_t<double, py::array::c_style |
                                            py::array::forcecast> a)
{
    if (a.ndim() != 2) {
        throw std::invalid_argument("expected a 2-D array");
    }
    num::rmat m(a.shape(0), a.shape(1));
    std::memcpy(m.a.data(), a.data(), m.a.size() * sizeof(double));
    return m;
}

py::dict problem_dict(const fdtd_problem& p)
{
    py::dict d;
    d["algo"] = (int)p.prob.algo;
    d["num_levels"] = p.num_lvl;
    d["num_materials"] = p.num_materials;
    d["num_gridpoints"] = (uint64_t)p.mat_idx.size();
    d["num_timesteps"] = p.num_timesteps;
    d["dx_inv"] = p.dx_inv;
    d["fdtd_consts"] = to_np(p.fdtd_consts);
    d["qm_consts"] = to_np(p.qm_consts);
    d["qm_stride"] = p.qm_stride;
    d["zform_consts"] = to_np(p.zform_consts);
    d["zform_stride"] = p.zform_stride;
    d["mat_idx"] = to_np(p.mat_idx);
    d["e_init"] = to_np(p.e_init);
    d["h_init"] = to_np(p.h_init);
    d["d_init"] = to_np(p.d_init);
    std::vector<uint32_t> st, sx;
    for (const auto& s : p.sources) {
        st.push_back(s.type);
        sx.push_back(s.x_idx);
    }
    d["source_type"] = to_np(st);
    d["source_x"] = to_np(sx);
    d["source_data"] = to_np(p.source_data);
    py::list recs;
    for (const auto& r : p.records) {
        py::dict rd;
        rd["type"] = r.type;
        rd["row_idx"] = r.row_idx;
        rd["col_idx"] = r.col_idx;
        rd["is_complex"] = r.is_complex;
        rd["rows"] = r.rows;
        rd["cols"] = r.cols;
        rd["position_idx"] = r.position_idx;
        rd["interval_ratio"] = r.interval_ratio;
        rd["offset_real"] = r.offset_real;
        rd["offset_imag"] = r.offset_imag;
        recs.append(rd);
    }
    d["records"] = recs;
    d["scratch_size"] = p.scratch_size;
    d["material_ids"] = p.material_ids;
    return d;
}

solver_hip_fdtd* as_hip(const std::shared_ptr<solver>& s)
{
    auto* h = dynamic_cast<solver_hip_fdtd*>(s.get());
    if (!h) {
        throw std::invalid_argument("not a hip-fdtd solver");
    }
    return h;
}

py::dict info_dict(const mbh_info& i)
{
    py::dict d;
    d["form"] = i.form;
    d["temporal_block"] = i.temporal_block;
    d["tile_cells"] = i.tile_cells;
    d["owned_cells"] = i.owned_cells;
    d["num_tiles"] = i.num_tiles;
    d["total_cells"] = i.total_cells;
    d["launches"] = i.launches;
    d["state_bytes"] = i.state_bytes;
    d["graph_launches"] = i.graph_launches;
    d["step_order"] = i.step_order;
    d["mfma_tiles"] = i.mfma_tiles;
    d["flags"] = i.flags;
    return d;
}

mbh_options make_options(uint32_t form, uint32_t tb, uint32_t tile,
                         uint32_t graph, int dev)
{
    mbh_options o;
    mbh_default_options(&o);
    o.form = form;
    o.temporal_block = tb;
    o.tile_cells = tile;
    o.graph_launches = graph;
    o.device = dev;
    return o;
}

} /* namespace */

PYBIND11_MODULE(_mbsolve, m)
{
    m.doc() = "mbsolve API with MI355X (gfx950) FDTD Maxwell-Bloch solvers";

    m.attr("HBAR") = HBAR;
    m.attr("PI") = PI;
    m.attr("MU0") = MU0;
    m.attr("EPS0") = EPS0;
    m.attr("E0") = E0;
    m.attr("KB") = KB;

    py::class_<qm_operator, std::shared_ptr<qm_operator> >(m, "qm_operator")
        .def(py::init<const std::vector<real>&,
                      const std::vector<std::complex<real> >&>(),
             py::arg("main_diagonal"),
             py::arg("off_diagonal") = std::vector<std::complex<real> >())
        .def("get_num_levels", &qm_operator::get_num_levels)
        .def("get_main_diagonal", &qm_operator::get_main_diagonal)
        .def("get_off_diagonal", &qm_operator::get_off_diagonal);

    py::class_<qm_superop, std::shared_ptr<qm_superop> >(m, "qm_superop")
        .def(py::init<unsigned int>())
        .def("action", &qm_superop::action)
        .def("__call__", &qm_superop::operator())
        .def("get_num_levels", &qm_superop::get_num_levels);

    py::class_<qm_lindblad_relaxation, qm_superop,
               std::shared_ptr<qm_lindblad_relaxation> >(
        m, "qm_lindblad_relaxation")
        .def(py::init<const std::vector<std::vector<real> >&,
                      const std::vector<real>&>())
        .def("get_scattering_rates",
             &qm_lindblad_relaxation::get_scattering_rates)
        .def("get_dephasing_rates",
             &qm_lindblad_relaxation::get_dephasing_rates);

    py::class_<qm_description, std::shared_ptr<qm_description> >(
        m, "qm_description")
        .def(py::init<real, const qm_operator&, const qm_operator&,
                      std::shared_ptr<qm_superop> >())
        .def("get_num_levels", &qm_description::get_num_levels)
        .def("get_carrier_density", &qm_description::get_carrier_density)
        .def("get_hamiltonian", &qm_description::get_hamiltonian)
        .def("get_dipole_operator", &qm_description::get_dipole_operator)
        .def("get_relaxation_superop",
             &qm_description::get_relaxation_superop);

    py::class_<qm_desc_2lvl, qm_description, std::shared_ptr<qm_desc_2lvl> >(
        m, "qm_desc_2lvl")
        .def(py::init<real, real, real, real, real, real>(),
             py::arg("carrier_density"), py::arg("transition_freq"),
             py::arg("dipole_moment"), py::arg("scattering_rate"),
             py::arg("dephasing_rate"),
             py::arg("equilibrium_inversion") = -1.0)
        .def("get_transition_freq", &qm_desc_2lvl::get_transition_freq)
        .def("get_dipole_moment", &qm_desc_2lvl::get_dipole_moment)
        .def("get_scattering_rate", &qm_desc_2lvl::get_scattering_rate)
        .def("get_dephasing_rate", &qm_desc_2lvl::get_dephasing_rate)
        .def("get_equilibrium_inversion",
             &qm_desc_2lvl::get_equilibrium_inversion);

    py::class_<material, std::shared_ptr<material> >(m, "material")
        .def(py::init<const std::string&, std::shared_ptr<qm_description>,
                      real, real, real, real>(),
             py::arg("id"), py::arg("qm") = nullptr,
             py::arg("rel_permittivity") = 1.0,
             py::arg("overlap_factor") = 1.0, py::arg("losses") = 0.0,
             py::arg("rel_permeability") = 1.0)
        .def("get_id", &material::get_id)
        .def("get_qm", &material::get_qm)
        .def("get_rel_permittivity", &material::get_rel_permittivity)
        .def("get_rel_permeability", &material::get_rel_permeability)
        .def("get_losses", &material::get_losses)
        .def("get_overlap_factor", &material::get_overlap_factor)
        .def_static("add_to_library", &material::add_to_library)
        .def_static("get_from_library", &material::get_from_library)
        .def_static("clear_library", &material::clear_library);

    py::class_<bc_field, std::shared_ptr<bc_field> >(m, "bc_field");
    py::class_<bc_field_reflectivity, bc_field,
               std::shared_ptr<bc_field_reflectivity> >(
        m, "bc_field_reflectivity")
        .def(py::init<real, real>(), py::arg("refl_left") = 1.0,
             py::arg("refl_right") = 1.0)
        .def("get_refl_left", &bc_field_reflectivity::get_refl_left)
        .def("get_refl_right", &bc_field_reflectivity::get_refl_right);

    py::class_<region, std::shared_ptr<region> >(m, "region")
        .def(py::init<const std::string&, std::shared_ptr<material>, real,
                      real>())
        .def("get_name", &region::get_name)
        .def("get_length", &region::get_length)
        .def("get_x_start", &region::get_x_start)
        .def("get_x_end", &region::get_x_end)
        .def("get_material", &region::get_material);

    py::class_<device, std::shared_ptr<device> >(m, "device")
        .def(py::init<const std::string&, std::shared_ptr<bc_field> >(),
             py::arg("name"),
             py::arg("field_boundary") =
                 std::make_shared<bc_field_reflectivity>(1.0, 1.0))
        .def("add_region", &device::add_region)
        .def("get_regions", &device::get_regions)
        .def("get_used_materials", &device::get_used_materials)
        .def("get_name", &device::get_name)
        .def("get_length", &device::get_length)
        .def("get_minimum_permittivity", &device::get_minimum_permittivity)
        .def("get_bc_field", &device::get_bc_field)
        .def("set_bc_field", &device::set_bc_field);

    py::class_<source, std::shared_ptr<source> > src(m, "source");
    py::enum_<source::type>(src, "type")
        .value("hard_source", source::hard_source)
        .value("soft_source", source::soft_source)
        .export_values();
    src.def("get_value", &source::get_value, py::arg("t"),
            py::arg("current_value") = 0.0)
        .def("calc_value", &source::calc_value)
        .def("get_name", &source::get_name)
        .def("get_position", &source::get_position)
        .def("get_type", &source::get_type);
    py::class_<sech_pulse, source, std::shared_ptr<sech_pulse> >(m,
                                                                 "sech_pulse")
        .def(py::init<const std::string&, real, source::type, real, real,
                      real, real, real>(),
             py::arg("name"), py::arg("position"), py::arg("source_type"),
             py::arg("ampl"), py::arg("freq"), py::arg("phase"),
             py::arg("beta"), py::arg("phase_sin") = 0.0);
    py::class_<gaussian_pulse, source, std::shared_ptr<gaussian_pulse> >(
        m, "gaussian_pulse")
        .def(py::init<const std::string&, real, source::type, real, real,
                      real, real>(),
             py::arg("name"), py::arg("position"), py::arg("source_type"),
             py::arg("ampl"), py::arg("freq"), py::arg("phase"),
             py::arg("tau"));
    py::class_<cw_source, source, std::shared_ptr<cw_source> >(m,
                                                               "cw_source")
        .def(py::init<const std::string&, real, source::type, real, real,
                      real>(),
             py::arg("name"), py::arg("position"), py::arg("source_type"),
             py::arg("ampl"), py::arg("freq"), py::arg("phase") = 0.0);

    py::class_<ic_density, std::shared_ptr<ic_density> >(m, "ic_density")
        .def("initialize", &ic_density::initialize);
    py::class_<ic_density_const, ic_density,
               std::shared_ptr<ic_density_const> >(m, "ic_density_const")
        .def(py::init<const qm_operator&>());
    py::class_<ic_field, std::shared_ptr<ic_field> >(m, "ic_field")
        .def("initialize", &ic_field::initialize);
    py::class_<ic_field_const, ic_field, std::shared_ptr<ic_field_const> >(
        m, "ic_field_const")
        .def(py::init<real>());
    py::class_<ic_field_random, ic_field, std::shared_ptr<ic_field_random> >(
        m, "ic_field_random")
        .def(py::init<real, real, real, unsigned int>(), py::arg("mean") = 0.0,
             py::arg("stddev") = 1.0, py::arg("amplitude") = 1e-15,
             py::arg("seed") = 0u);

    py::class_<record, std::shared_ptr<record> > rec(m, "record");
    py::enum_<record::type>(rec, "type")
        .value("electric", record::electric)
        .value("polar_dt", record::polar_dt)
        .value("magnetic", record::magnetic)
        .value("density", record::density)
        .value("inversion", record::inversion)
        .export_values();
    rec.def(py::init<const std::string&, real, real>(), py::arg("name"),
            py::arg("interval") = 0.0, py::arg("position") = -1.0)
        .def(py::init<const std::string&, record::type, unsigned int,
                      unsigned int, real, real>(),
             py::arg("name"), py::arg("record_type"), py::arg("row_idx"),
             py::arg("col_idx"), py::arg("interval") = 0.0,
             py::arg("position") = -1.0)
        .def("get_name", &record::get_name)
        .def("get_type", &record::get_type)
        .def("get_row", &record::get_row)
        .def("get_col", &record::get_col)
        .def("get_position", &record::get_position)
        .def("get_interval", &record::get_interval)
        .def("is_complex", &record::is_complex);

    py::class_<scenario, std::shared_ptr<scenario> >(m, "scenario")
        .def(py::init<const std::string&, unsigned int, real,
                      const qm_operator&, unsigned int>(),
             py::arg("name"), py::arg("num_gridpoints"), py::arg("endtime"),
             py::arg("rho_init"), py::arg("num_timesteps") = 2)
        .def(py::init<const std::string&, unsigned int, real,
                      std::shared_ptr<ic_density>, std::shared_ptr<ic_field>,
                      std::shared_ptr<ic_field>, unsigned int>(),
             py::arg("name"), py::arg("num_gridpoints"), py::arg("endtime"),
             py::arg("density_init"),
             py::arg("electric_init") = std::make_shared<ic_field_random>(),
             py::arg("magnetic_init") = std::make_shared<ic_field_const>(0),
             py::arg("num_timesteps") = 2)
        .def("add_record", &scenario::add_record)
        .def("get_records", &scenario::get_records)
        .def("add_source", &scenario::add_source)
        .def("get_sources", &scenario::get_sources)
        .def("get_name", &scenario::get_name)
        .def("get_num_timesteps", &scenario::get_num_timesteps)
        .def("set_num_timesteps", &scenario::set_num_timesteps)
        .def("get_num_gridpoints", &scenario::get_num_gridpoints)
        .def("set_num_gridpoints", &scenario::set_num_gridpoints)
        .def("get_timestep_size", &scenario::get_timestep_size)
        .def("set_timestep_size", &scenario::set_timestep_size)
        .def("get_gridpoint_size", &scenario::get_gridpoint_size)
        .def("set_gridpoint_size", &scenario::set_gridpoint_size)
        .def("get_endtime", &scenario::get_endtime)
        .def("set_endtime", &scenario::set_endtime)
        .def("get_ic_density", &scenario::get_ic_density)
        .def("set_ic_density", &scenario::set_ic_density)
        .def("get_ic_electric", &scenario::get_ic_electric)
        .def("set_ic_electric", &scenario::set_ic_electric)
        .def("get_ic_magnetic", &scenario::get_ic_magnetic)
        .def("set_ic_magnetic", &scenario::set_ic_magnetic);

    py::class_<result, std::shared_ptr<result> >(m, "result")
        .def(py::init<const std::string&, unsigned int, unsigned int, bool>(),
             py::arg("name"), py::arg("cols"), py::arg("rows"),
             py::arg("is_complex") = false)
        .def("set_data",
             [](result& r, py::array_t<double, py::array::c_style |
                                                   py::array::forcecast> re,
                py::object im) {
                 /* fills the real (and imaginary) part, row-major
                  * [rows][cols]: for results built outside a solver */
                 if ((size_t)re.size() != (size_t)r.get_count()) {
                     throw std::invalid_argument("result size mismatch");
                 }
                 std::copy(re.data(), re.data() + re.size(),
                           r.get_data_real_it(0, 0));
                 if (!im.is_none()) {
                     auto ia = im.cast<py::array_t<
                         double, py::array::c_style | py::array::forcecast> >();
                     if ((size_t)ia.size() != (size_t)r.get_count()) {
                         throw std::invalid_argument("result size mismatch");
                     }
                     std::copy(ia.data(), ia.data() + ia.size(),
                               r.get_data_imag_it(0, 0));
                 }
             },
             py::arg("real"), py::arg("imag") = py::none())
        .def("get_name", &result::get_name)
        .def("get_cols", &result::get_cols)
        .def("get_rows", &result::get_rows)
        .def("get_count", &result::get_count)
        .def("is_complex", &result::is_complex)
        .def("get_data_real",
             [](const result& r) {
                 py::array_t<double> a({ (size_t)r.get_rows(),
                                         (size_t)r.get_cols() });
                 std::memcpy(a.mutable_data(), r.get_data_real().data(),
                             r.get_data_real().size() * sizeof(double));
                 return a;
             })
        .def("get_data_imag", [](const result& r) {
            py::array_t<double> a({ (size_t)r.get_rows(),
                                    (size_t)r.get_cols() });
            std::memcpy(a.mutable_data(), r.get_data_imag().data(),
                        r.get_data_imag().size() * sizeof(double));
            return a;
        });

    py::class_<solver, std::shared_ptr<solver> >(m, "solver")
        .def_static("create_instance", &solver::create_instance)
        .def_static("get_avail_solvers", &solver::get_avail_solvers)
        .def("get_name", &solver::get_name)
        .def("get_results", &solver::get_results)
        .def("run", &solver::run, py::call_guard<py::gil_scoped_release>())
        /* HIP solver extensions */
        .def("run_steps",
             [](const std::shared_ptr<solver>& s, uint64_t a, uint64_t b) {
                 py::gil_scoped_release g;
                 as_hip(s)->run_steps(a, b);
             })
        .def("fetch_results",
             [](const std::shared_ptr<solver>& s) {
                 py::gil_scoped_release g;
                 as_hip(s)->fetch_results();
             })
        .def("reset",
             [](const std::shared_ptr<solver>& s) { as_hip(s)->reset(); })
        .def("elapsed_ms",
             [](const std::shared_ptr<solver>& s) {
                 return as_hip(s)->elapsed_ms();
             })
        .def("info",
             [](const std::shared_ptr<solver>& s) {
                 return info_dict(as_hip(s)->info());
             })
        .def("problem",
             [](const std::shared_ptr<solver>& s) {
                 return problem_dict(as_hip(s)->problem());
             })
        .def("fetch_state", [](const std::shared_ptr<solver>& s) {
            auto* h = as_hip(s);
            const auto& p = h->problem();
            const size_t nx = p.mat_idx.size();
            const size_t ds = nx ? p.d_init.size() / nx : 0;
            std::vector<real> e(nx), hh(nx + 1), pp(nx), d(nx * ds);
            h->fetch_state(e.data(), hh.data(), pp.data(),
                           ds ? d.data() : nullptr);
            py::dict out;
            out["e"] = to_np(e);
            out["h"] = to_np(hh);
            out["p"] = to_np(pp);
            out["d"] = to_np(d);
            return out;
        });

    py::class_<writer, std::shared_ptr<writer> >(m, "writer")
        .def_static("create_instance", &writer::create_instance)
        .def_static("get_avail_writers", &writer::get_avail_writers)
        .def("get_name", &writer::get_name)
        .def("get_extension", &writer::get_extension)
        .def("write", &writer::write);

    py::class_<solver_hip_batch, std::shared_ptr<solver_hip_batch> >(
        m, "solver_hip_batch")
        .def(py::init([](const std::string& name,
                         const std::vector<std::pair<
                             std::shared_ptr<device>,
                             std::shared_ptr<scenario> > >& runs,
                         uint32_t tb, uint32_t tile, uint32_t graph) {
                 std::vector<std::pair<std::shared_ptr<const device>,
                                       std::shared_ptr<scenario> > >
                     r;
                 for (auto& x : runs) {
                     r.emplace_back(x.first, x.second);
                 }
                 density_algo a;
                 unsigned int l;
                 uint32_t form;
                 if (!parse_hip_solver_name(name, a, l, form)) {
                     throw std::invalid_argument("Unknown solver " + name);
                 }
                 mbh_options o = hip_options_from_env(form);
                 if (tb) {
                     o.temporal_block = tb;
                 }
                 if (tile) {
                     o.tile_cells = tile;
                 }
                 if (graph) {
                     o.graph_launches = graph;
                 }
                 return std::make_shared<solver_hip_batch>(name, r, &o);
             }),
             py::arg("name"), py::arg("runs"), py::arg("temporal_block") = 0,
             py::arg("tile_cells") = 0, py::arg("graph_launches") = 0)
        .def("get_name", &solver_hip_batch::get_name)
        .def("num_runs", &solver_hip_batch::num_runs)
        .def("run", &solver_hip_batch::run,
             py::call_guard<py::gil_scoped_release>())
        .def("run_steps", &solver_hip_batch::run_steps,
             py::call_guard<py::gil_scoped_release>())
        .def("fetch_results", &solver_hip_batch::fetch_results,
             py::call_guard<py::gil_scoped_release>())
        .def("reset", &solver_hip_batch::reset)
        .def("elapsed_ms", &solver_hip_batch::elapsed_ms)
        .def("info",
             [](const solver_hip_batch& b) { return info_dict(b.info()); })
        .def("get_results", &solver_hip_batch::get_results)
        .def("problem", [](const solver_hip_batch& b, size_t r) {
            return problem_dict(b.problem(r));
        });

    m.def(
        "load_hip_solvers", []() { solver_hip_loader l; },
        "Registers the hip-fdtd-* solver names (solver_hip_loader).");

    m.def(
        "create_hip_solver",
        [](const std::string& name, std::shared_ptr<device> dev,
           std::shared_ptr<scenario> scen, uint32_t tb, uint32_t tile,
           uint32_t graph, int devnum) {
            density_algo a;
            unsigned int l;
            uint32_t form;
            if (!parse_hip_solver_name(name, a, l, form)) {
                throw std::invalid_argument("Unknown solver " + name);
            }
            mbh_options o = hip_options_from_env(form);
            if (tb) {
                o.temporal_block = tb;
            }
            if (tile) {
                o.tile_cells = tile;
            }
            if (graph) {
                o.graph_launches = graph;
            }
            if (devnum >= 0) {
                o.device = devnum;
            }
            return std::static_pointer_cast<solver>(
                std::make_shared<solver_hip_fdtd>(dev, scen, name, &o));
        },
        py::arg("name"), py::arg("dev"), py::arg("scen"),
        py::arg("temporal_block") = 0, py::arg("tile_cells") = 0,
        py::arg("graph_launches") = 0, py::arg("device") = -1,
        "Creates a hip-fdtd solver with explicit launch options.");

    m.def(
        "build_problem",
        [](const std::string& name, std::shared_ptr<device> dev,
           std::shared_ptr<scenario> scen) {
            density_algo a;
            unsigned int l;
            uint32_t form;
            if (!parse_hip_solver_name(name, a, l, form)) {
                throw std::invalid_argument("Unknown solver " + name);
            }
            fdtd_problem p;
            const bool zf = (a == density_algo::cvr_rodr) &&
                form != MBH_FORM_FP64;
            build_fdtd_problem(dev, scen, a, l, zf, p);
            return problem_dict(p);
        },
        "Host set-up of a hip-fdtd solver without touching the GPU.");

    m.def(
        "init_fdtd_simulation",
        [](std::shared_ptr<device> dev, std::shared_ptr<scenario> scen,
           real courant) { init_fdtd_simulation(dev, scen, courant); },
        py::arg("dev"), py::arg("scen"), py::arg("courant") = 1.0);

    /* host numerics, exposed for verification */
    py::module_ nm = m.def_submodule("numerics", "host FP64 numerics");
    nm.def("expm", [](py::array_t<double, py::array::c_style |
                                              py::array::forcecast> a) {
        return mat_to_np(num::expm(np_to_mat(a)));
    });
    nm.def("skew_decompose",
           [](py::array_t<double, py::array::c_style | py::array::forcecast>
                  a) {
               auto sb = num::skew_decompose(np_to_mat(a));
               return py::make_tuple(mat_to_np(sb.Q), to_np(sb.theta));
           });
    nm.def("solve_full_pivot",
           [](py::array_t<double, py::array::c_style | py::array::forcecast>
                  a,
              const std::vector<real>& b) {
               return to_np(num::solve_full_pivot(np_to_mat(a), b));
           });
    nm.def("cayley_constants",
           [](std::shared_ptr<qm_description> qm, real dt, unsigned int n) {
               return to_np(cayley_constants(qm, dt, n));
           });
    nm.def("rodr_constants",
           [](std::shared_ptr<qm_description> qm, real dt, unsigned int n) {
               std::vector<real> b, z;
               rodr_constants(qm, dt, n, b, &z);
               return py::make_tuple(to_np(b), to_np(z));
           });
    nm.def("cv_matrices", [](std::shared_ptr<qm_description> qm) {
        cv_representation c(qm);
        py::dict d;
        d["hamiltonian"] = mat_to_np(c.get_hamiltonian());
        d["dipole_operator"] = mat_to_np(c.get_dipole_operator());
        d["dipole_operator_vec"] = to_np(c.get_dipole_operator_vec());
        d["relaxation_superop"] = mat_to_np(c.get_relaxation_superop());
        d["equilibrium_vec"] = to_np(c.get_equilibrium_vec());
        return d;
    });
    nm.def("coherence_vector",
           [](const qm_operator& op, unsigned int n) {
               return to_np(cv_representation::coherence_vector(op, n));
           });
    nm.def("calc_population",
           [](const std::vector<real>& d, unsigned int n, unsigned int mm) {
               return cv_representation::calc_population(d, n, mm);
           });

    /* C ABI facts */
    m.def("abi_version", []() { return MBH_ABI_VERSION; });
    m.def("library_version", []() { return std::string(mbh_version()); });
    m.def("device_count", []() {
        int n = 0;
        mbh_device_count(&n);
        return n;
    });
}
