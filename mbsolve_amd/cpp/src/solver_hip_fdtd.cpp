/*
 * HIP solver classes (see solver_hip_fdtd.hpp).
 */
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <mbsolve/solver-hip/solver_hip_fdtd.hpp>

namespace mbsolve {

namespace {

void check(int rc, const mbh_ctx* ctx, const char* what)
{
    if (rc != MBH_OK) {
        throw std::runtime_error(std::string(what) + ": " +
                                 mbh_last_error(ctx));
    }
}

uint32_t env_u32(const char* name, uint32_t dflt)
{
    const char* v = std::getenv(name);
    if (!v || !*v) {
        return dflt;
    }
    return (uint32_t)std::strtoul(v, nullptr, 10);
}

void copy_results(mbh_ctx* ctx, uint32_t run, const fdtd_problem& p)
{
    if (p.scratch_size == 0) {
        return;
    }
    std::vector<real> scratch(p.scratch_size);
    check(mbh_fetch_records(ctx, run, scratch.data(), p.scratch_size), ctx,
          "mbh_fetch_records");
    for (const auto& e : p.copy_list) {
        const real* src = scratch.data() + e.dev.offset_real;
        std::copy(src, src + e.size(), e.res->get_data_real_it(0, 0));
        if (e.dev.is_complex) {
            const real* si = scratch.data() + e.dev.offset_imag;
            std::copy(si, si + e.size(), e.res->get_data_imag_it(0, 0));
        }
    }
}

} /* namespace */

bool
parse_hip_solver_name(const std::string& name_in, density_algo& algo,
                      unsigned int& num_lvl, uint32_t& form)
{
    const std::string pre = "hip-fdtd-";
    if (name_in.compare(0, pre.size(), pre) != 0) {
        return false;
    }
    std::string rest = name_in.substr(pre.size());
    /* "red-": the step order of solver_cpu_fdtd_red (FP64 form only) */
    const bool red = rest.compare(0, 4, "red-") == 0;
    if (red) {
        rest = rest.substr(4);
    }
    form = MBH_FORM_FP64;
    auto strip = [&rest](const std::string& suf) {
        if (rest.size() > suf.size() &&
            rest.compare(rest.size() - suf.size(), suf.size(), suf) == 0) {
            rest = rest.substr(0, rest.size() - suf.size());
            return true;
        }
        return false;
    };
    if (strip("-mixed")) {
        form = MBH_FORM_MIXED;
    } else if (strip("-zform")) {
        form = MBH_FORM_ZFORM64;
    } else if (strip("-mfma64")) {
        form = MBH_FORM_MFMA64;
    } else if (strip("-mfma")) {
        form = MBH_FORM_MFMA;
    }
    const bool mfma = (form == MBH_FORM_MFMA || form == MBH_FORM_MFMA64);
    if (red && form != MBH_FORM_FP64) {
        return false;
    }
    if (rest == "noop") {
        algo = density_algo::noop;
        num_lvl = 0;
        return form == MBH_FORM_FP64 || form == MBH_FORM_MIXED;
    }
    const size_t p = rest.find("lvl-");
    if (p == std::string::npos || p == 0) {
        return false;
    }
    const std::string lv = rest.substr(0, p), alg = rest.substr(p + 4);
    if (lv.size() != 1 || lv[0] < '2' || lv[0] > '6') {
        return false;
    }
    num_lvl = lv[0] - '0';
    if (alg == "rodr") {
        algo = density_algo::cvr_rodr;
        return !mfma || num_lvl >= 3;
    }
    if (alg == "reg-cayley") {
        algo = density_algo::reg_cayley;
        return !mfma && (num_lvl == 2 || form != MBH_FORM_ZFORM64);
    }
    return false;
}

uint32_t
hip_solver_order(const std::string& name)
{
    return name.compare(0, 13, "hip-fdtd-red-") == 0 ? MBH_ORDER_RED
                                                     : MBH_ORDER_PLAIN;
}

mbh_options
hip_options_from_env(uint32_t form)
{
    mbh_options o;
    mbh_default_options(&o);
    o.form = form;
    o.temporal_block = env_u32("MBSOLVE_HIP_TEMPORAL_BLOCK", 0);
    o.tile_cells = env_u32("MBSOLVE_HIP_TILE_CELLS", 0);
    o.graph_launches = env_u32("MBSOLVE_HIP_GRAPH_LAUNCHES", 0);
    const char* d = std::getenv("MBSOLVE_HIP_DEVICE");
    if (d && *d) {
        o.device = std::atoi(d);
    }
    return o;
}

/* ---- single run ---------------------------------------------------------- */
solver_hip_fdtd::solver_hip_fdtd(std::shared_ptr<const device> dev,
                                 std::shared_ptr<scenario> scen,
                                 const std::string& name,
                                 const mbh_options* opt)
  : solver(name, dev, scen)
{
    density_algo algo;
    unsigned int lvl;
    if (!parse_hip_solver_name(name, algo, lvl, m_form)) {
        throw std::invalid_argument("Unknown solver " + name);
    }
    const bool zf = (algo == density_algo::cvr_rodr) && m_form != MBH_FORM_FP64;
    build_fdtd_problem(dev, scen, algo, lvl, zf, m_problem);
    for (const auto& e : m_problem.copy_list) {
        m_results.push_back(e.res);
    }
    mbh_options o = opt ? *opt : hip_options_from_env(m_form);
    o.form = m_form;
    o.step_order = hip_solver_order(name);
    check(mbh_create(&m_problem.prob, 1, &o, &m_ctx), nullptr, "mbh_create");
}

solver_hip_fdtd::~solver_hip_fdtd() { mbh_destroy(m_ctx); }

void
solver_hip_fdtd::run_steps(uint64_t n_begin, uint64_t n_end) const
{
    check(mbh_run(m_ctx, n_begin, n_end), m_ctx, "mbh_run");
}

void
solver_hip_fdtd::reset() const
{
    check(mbh_reset(m_ctx), m_ctx, "mbh_reset");
}

void
solver_hip_fdtd::fetch_results() const
{
    copy_results(m_ctx, 0, m_problem);
}

void
solver_hip_fdtd::run() const
{
    run_steps(0, m_problem.num_timesteps);
    fetch_results();
}

double
solver_hip_fdtd::elapsed_ms() const
{
    double ms = 0.0;
    check(mbh_elapsed_ms(m_ctx, &ms), m_ctx, "mbh_elapsed_ms");
    return ms;
}

mbh_info
solver_hip_fdtd::info() const
{
    mbh_info i;
    check(mbh_get_info(m_ctx, &i), m_ctx, "mbh_get_info");
    return i;
}

void
solver_hip_fdtd::fetch_state(real* e, real* h, real* p, real* d) const
{
    check(mbh_fetch_state(m_ctx, 0, e, h, p, d), m_ctx, "mbh_fetch_state");
}

/* ---- batch ---------------------------------------------------------------- */
solver_hip_batch::solver_hip_batch(
    const std::string& name,
    const std::vector<std::pair<std::shared_ptr<const device>,
                                std::shared_ptr<scenario> > >& runs,
    const mbh_options* opt)
  : m_name(name), m_runs(runs)
{
    density_algo algo;
    unsigned int lvl;
    uint32_t form;
    if (!parse_hip_solver_name(name, algo, lvl, form)) {
        throw std::invalid_argument("Unknown solver " + name);
    }
    if (runs.empty()) {
        throw std::invalid_argument("Empty batch.");
    }
    const bool zf = (algo == density_algo::cvr_rodr) && form != MBH_FORM_FP64;
    m_problems.resize(runs.size());
    for (size_t r = 0; r < runs.size(); r++) {
        build_fdtd_problem(runs[r].first, runs[r].second, algo, lvl, zf,
                           m_problems[r]);
    }
    std::vector<mbh_problem> probs;
    for (auto& p : m_problems) {
        p.bind();
        probs.push_back(p.prob);
    }
    mbh_options o = opt ? *opt : hip_options_from_env(form);
    o.form = form;
    o.step_order = hip_solver_order(name);
    check(mbh_create(probs.data(), (uint32_t)probs.size(), &o, &m_ctx),
          nullptr, "mbh_create");
}

solver_hip_batch::~solver_hip_batch() { mbh_destroy(m_ctx); }

void
solver_hip_batch::run_steps(uint64_t n_begin, uint64_t n_end) const
{
    check(mbh_run(m_ctx, n_begin, n_end), m_ctx, "mbh_run");
}

void
solver_hip_batch::reset() const
{
    check(mbh_reset(m_ctx), m_ctx, "mbh_reset");
}

void
solver_hip_batch::fetch_results() const
{
    for (size_t r = 0; r < m_problems.size(); r++) {
        copy_results(m_ctx, (uint32_t)r, m_problems[r]);
    }
}

void
solver_hip_batch::run() const
{
    run_steps(0, m_problems[0].num_timesteps);
    fetch_results();
}

double
solver_hip_batch::elapsed_ms() const
{
    double ms = 0.0;
    check(mbh_elapsed_ms(m_ctx, &ms), m_ctx, "mbh_elapsed_ms");
    return ms;
}

mbh_info
solver_hip_batch::info() const
{
    mbh_info i;
    check(mbh_get_info(m_ctx, &i), m_ctx, "mbh_get_info");
    return i;
}

std::vector<std::shared_ptr<result> >
solver_hip_batch::get_results(size_t r) const
{
    std::vector<std::shared_ptr<result> > out;
    for (const auto& e : m_problems.at(r).copy_list) {
        out.push_back(e.res);
    }
    return out;
}

/* ---- loader --------------------------------------------------------------- */
solver_hip_loader::solver_hip_loader()
{
    static std::once_flag once;
    std::call_once(once, []() {
        std::vector<std::string> names;
        for (const char* f : { "", "-mixed" }) {
            names.push_back(std::string("hip-fdtd-noop") + f);
        }
        for (unsigned int n = 2; n <= 6; n++) {
            for (const char* f : { "", "-mixed", "-zform", "-mfma",
                                   "-mfma64" }) {
                if (n < 3 && std::string(f).rfind("-mfma", 0) == 0) {
                    continue;
                }
                names.push_back("hip-fdtd-" + std::to_string(n) + "lvl-rodr" +
                                f);
            }
        }
        for (const char* f : { "", "-mixed", "-zform" }) {
            names.push_back(std::string("hip-fdtd-2lvl-reg-cayley") + f);
        }
        for (unsigned int n = 3; n <= 6; n++) {
            for (const char* f : { "", "-mixed" }) {
                names.push_back("hip-fdtd-" + std::to_string(n) +
                                "lvl-reg-cayley" + f);
            }
        }
        /* step order of solver_cpu_fdtd_red (solver_cpu_loader.cpp:38-40,
         * 54-64, 78-88) */
        names.push_back("hip-fdtd-red-noop");
        for (unsigned int n = 2; n <= 6; n++) {
            names.push_back("hip-fdtd-red-" + std::to_string(n) + "lvl-rodr");
            names.push_back("hip-fdtd-red-" + std::to_string(n) +
                            "lvl-reg-cayley");
        }
        for (const auto& nm : names) {
            solver::register_bootstrap(
                nm, [nm](std::shared_ptr<const device> dev,
                         std::shared_ptr<scenario> scen) {
                    return std::make_shared<solver_hip_fdtd>(dev, scen, nm);
                });
        }
    });
}
}
