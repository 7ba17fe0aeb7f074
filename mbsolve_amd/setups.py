"""Scenario recipes as plain-data specs, and a builder that turns a spec into
mbsolve API objects (device, scenario).

A spec is a dict::

    {"name": str, "algo": "reg-cayley" | "cvr-rodr" | "noop",
     "num_levels": N,
     "materials": {id: {"rel_permittivity", "overlap_factor", "losses",
                        "rel_permeability", "qm": qm-spec or None}},
     "regions": [(name, material_id, x_start, x_end), ...],
     "num_gridpoints": n_x, "endtime": T, "num_timesteps": n_t (0-D only),
     "rho_init": (main_diagonal, off_diagonal),
     "e_init": E0, "h_init": H0,
     "sources": [{"kind": "sech"|"gauss"|"cw", "name", "position", "hard",
                  "ampl", "freq", "phase", "beta", "phase_sin", "tau"}],
     "records": [{"name", "interval", "position", ["type", "row", "col"]}]}

qm-spec is {"type": "2lvl", carrier_density, transition_freq, dipole_moment,
scattering_rate, dephasing_rate, equilibrium_inversion} (qm_desc_2lvl) or
{"num_levels", "carrier_density", "h_diag", "h_off", "mu_diag", "mu_off",
"scattering", "pure_dephasing"} (qm_description with Lindblad relaxation).

Recipes: the two-level self-induced-transparency run (sit_2lvl), the
three-level V system (v3_rodr), the six-level gain medium (qcl6), their
batched sweeps, and the setups of the reference's command-line tool
(ziolkowski1995, song2005, marskar2011, tzenov2018_cpml; mbsolve-tool.cpp).
"""
import math

import numpy as np

HBAR = 1.05457266e-34
PI = math.atan(1.0) * 4.0
E0 = 1.60217733e-19
KB = 1.38064852e-23
C0 = 1.0 / math.sqrt(PI * 4e-7 * 8.854187817e-12)

SOLVER_ALGO = {"reg-cayley": "reg-cayley", "rodr": "cvr-rodr",
               "noop": "noop"}


# ------------------------------------------------------------- builder -----
def build(spec, clear_library=True):
    """Creates (device, scenario) from a spec with the mbsolve API."""
    from . import lib as mb
    if clear_library:
        mb.material.clear_library()
    mats = {}
    for mid, m in spec["materials"].items():
        if not clear_library:
            # runs of one batch share materials of the same name (e.g. the
            # vacuum); the first definition stays in the library
            try:
                mats[mid] = mb.material.get_from_library(mid)
                continue
            except Exception:
                pass
        qm = _make_qm(mb, m.get("qm"))
        mat = mb.material(mid, qm, m.get("rel_permittivity", 1.0),
                          m.get("overlap_factor", 1.0),
                          m.get("losses", 0.0),
                          m.get("rel_permeability", 1.0))
        mb.material.add_to_library(mat)
        mats[mid] = mat
    refl = spec.get("reflectivity")
    dev = (mb.device(spec.get("name", "device"),
                     mb.bc_field_reflectivity(*refl)) if refl else
           mb.device(spec.get("name", "device")))
    for name, mid, x0, x1 in spec["regions"]:
        dev.add_region(mb.region(name, mats[mid], x0, x1))
    diag, off = spec["rho_init"]
    # "e_init_random": seed of the reference's default random initial field
    # (ic_field_random: N(0, 1) x 1e-15), seeded for reproducibility
    e_ic = (mb.ic_field_random(seed=int(spec["e_init_random"]))
            if "e_init_random" in spec else
            mb.ic_field_const(spec.get("e_init", 0.0)))
    scen = mb.scenario(spec.get("name", "scenario"), spec["num_gridpoints"],
                       spec["endtime"],
                       mb.ic_density_const(mb.qm_operator(list(diag),
                                                          list(off))),
                       e_ic,
                       mb.ic_field_const(spec.get("h_init", 0.0)),
                       spec.get("num_timesteps", 2))
    for s in spec.get("sources", []):
        st = mb.source.hard_source if s.get("hard", True) \
            else mb.source.soft_source
        if s["kind"] == "sech":
            src = mb.sech_pulse(s["name"], s["position"], st, s["ampl"],
                                s["freq"], s["phase"], s["beta"],
                                s.get("phase_sin", 0.0))
        elif s["kind"] == "gauss":
            src = mb.gaussian_pulse(s["name"], s["position"], st, s["ampl"],
                                    s["freq"], s["phase"], s["tau"])
        elif s["kind"] == "cw":
            src = mb.cw_source(s["name"], s["position"], st, s["ampl"],
                               s["freq"], s.get("phase", 0.0))
        else:
            raise ValueError("unknown source kind %r" % s["kind"])
        scen.add_source(src)
    for r in spec.get("records", []):
        if "type" in r:
            scen.add_record(mb.record(r["name"], mb.record.type(r["type"]),
                                      r.get("row", 1), r.get("col", 1),
                                      r.get("interval", 0.0),
                                      r.get("position", -1.0)))
        else:
            scen.add_record(mb.record(r["name"], r.get("interval", 0.0),
                                      r.get("position", -1.0)))
    return dev, scen


def _make_qm(mb, q):
    if q is None:
        return None
    if q.get("type") == "2lvl":
        return mb.qm_desc_2lvl(q["carrier_density"], q["transition_freq"],
                               q["dipole_moment"], q["scattering_rate"],
                               q["dephasing_rate"],
                               q.get("equilibrium_inversion", -1.0))
    H = mb.qm_operator(list(q["h_diag"]), [complex(x) for x in q["h_off"]])
    u = mb.qm_operator(list(q["mu_diag"]), [complex(x) for x in q["mu_off"]])
    sop = mb.qm_lindblad_relaxation([list(map(float, r))
                                     for r in q["scattering"]],
                                    list(map(float, q["pure_dephasing"])))
    return mb.qm_description(q["carrier_density"], H, u, sop)


def solver_name(spec, form="", order="plain"):
    """hip-fdtd-* name for the spec's algorithm ('' / 'mixed' / 'zform' /
    'mfma' / 'mfma64'); order 'red' gives the hip-fdtd-red-* name."""
    a = spec["algo"]
    pre = "hip-fdtd-red-" if order == "red" else "hip-fdtd-"
    base = (pre + "noop" if a == "noop" else
            pre + "%dlvl-%s" % (spec["num_levels"],
                                "rodr" if a == "cvr-rodr" else a))
    return base + ("-" + form if form else "")


# ------------------------------------------------------------- recipes -----
def sech_area_amplitude(area, beta, dipole):
    """Amplitude of a sech pulse of pulse area `area` (A = area hbar beta /
    (pi d), so area 2 pi gives the 2 hbar beta / d of self-induced
    transparency)."""
    return area * HBAR * beta / (PI * dipole)


def sit_2lvl(num_gridpoints=8192, endtime=1.5e-12, area=2 * PI, t2=1e-9,
             t1=1e-9, num_records=256, length=150e-6, phase=7.0, beta=1e13,
             name="sit2"):
    """Two-level self-induced transparency: 15 um vacuum, 120 um two-level
    medium (2 pi x 2e14 rad/s, 1e-29 C m, 1e24 m^-3, T1 = T2 = 1 ns, ground
    state, eps_r = 1), 15 um vacuum; sech pulse at x = 0 with resonant
    carrier, beta = 1e13 1/s, t0 = phase / beta; inversion and field
    recorded at `num_records` equispaced times over the run."""
    d = 1e-29
    margin = length * 0.1
    qm = {"type": "2lvl", "carrier_density": 1e24,
          "transition_freq": 2 * PI * 2e14, "dipole_moment": d / E0,
          "scattering_rate": 1.0 / t1, "dephasing_rate": 1.0 / t2,
          "equilibrium_inversion": -1.0}
    interval = endtime / (num_records - 1) if num_records > 1 else 0.0
    return {
        "name": name, "algo": "reg-cayley", "num_levels": 2,
        "materials": {"Vacuum": {}, "AR_" + name: {"qm": qm}},
        "regions": [("Vacuum left", "Vacuum", 0.0, margin),
                    ("Active region", "AR_" + name, margin, length - margin),
                    ("Vacuum right", "Vacuum", length - margin, length)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([1.0, 0.0], []),
        "sources": [{"kind": "sech", "name": "sech", "position": 0.0,
                     "hard": True,
                     "ampl": sech_area_amplitude(area, beta, d),
                     "freq": 2e14, "phase": phase, "beta": beta}],
        "records": [{"name": "inv12", "interval": interval},
                    {"name": "e", "interval": interval}],
    }


def v3_rodr(num_gridpoints=65536, endtime=1.5e-12, seed=1, num_records=256,
            length=150e-6, algo="cvr-rodr", name="v3"):
    """Three-level V system on the Rodrigues propagator: levels 0,
    hbar 2 pi 2e14, hbar 2 pi 2.2e14; dipoles (0,1) = 1e-29 C m,
    (0,2) = 0.8e-29 C m; scattering 1e10 1/s from levels 1, 2 to 0
    (T1 = 100 ps); pure dephasing such that every coherence decays at
    1e12 1/s (T2 = 1 ps).  Two soft sech sources at x = 0, one per carrier,
    with pulse areas drawn from U[0.5 pi, 3 pi] (seeded); inversion
    rho11 - rho00 and field recorded at `num_records` times."""
    rng = np.random.default_rng(seed)
    areas = rng.uniform(0.5 * PI, 3 * PI, size=2)
    d1, d2 = 1e-29, 0.8e-29
    rate = 1e10
    scat = [[0.0, rate, rate], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    # total dephasing = pure + (|s_mm| + |s_nn|)/2; out-rates: level 0: 0,
    # levels 1 and 2: 1e10 each
    pure = [1e12 - 0.5 * rate, 1e12 - 0.5 * rate, 1e12 - rate]
    qm = {"num_levels": 3, "carrier_density": 1e24,
          "h_diag": [0.0, HBAR * 2 * PI * 2e14, HBAR * 2 * PI * 2.2e14],
          "h_off": [0.0, 0.0, 0.0], "mu_diag": [0.0, 0.0, 0.0],
          "mu_off": [-d1, -d2, 0.0], "scattering": scat,
          "pure_dephasing": pure}
    beta = 1e13
    interval = endtime / (num_records - 1) if num_records > 1 else 0.0
    margin = length * 0.1
    return {
        "name": name, "algo": algo, "num_levels": 3,
        "materials": {"Vacuum": {}, "AR_" + name: {"qm": qm}},
        "regions": [("Vacuum left", "Vacuum", 0.0, margin),
                    ("Active region", "AR_" + name, margin, length - margin),
                    ("Vacuum right", "Vacuum", length - margin, length)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([1.0, 0.0, 0.0], []),
        "sources": [
            {"kind": "sech", "name": "sech1", "position": 0.0, "hard": False,
             "ampl": sech_area_amplitude(areas[0], beta, d1), "freq": 2e14,
             "phase": 7.0, "beta": beta},
            {"kind": "sech", "name": "sech2", "position": 0.0, "hard": False,
             "ampl": sech_area_amplitude(areas[1], beta, d2), "freq": 2.2e14,
             "phase": 7.0, "beta": beta}],
        "records": [{"name": "inv12", "interval": interval},
                    {"name": "e", "interval": interval}],
        "areas": [float(a) for a in areas],
    }


def qcl6(num_gridpoints=1 << 20, num_timesteps=4000000, seed=1,
         bias=1.0, length=None, num_records=64, name="qcl6",
         algo="cvr-rodr", cw_ampl=1e3):
    """Six-level quantum-cascade-style gain medium (35-dim Bloch vector):
    seeded random real symmetric Hamiltonian with level spacings
    U[5, 30] meV (scaled by `bias`) and nearest-neighbour tunnel couplings
    U[0.5, 2] meV, dipoles -E0 U[1, 5] nm on the (n, n+1) and (n, n+2)
    transitions, scattering rates U[1e11, 1e13] 1/s between all levels and
    pure dephasing U[1e11, 1e13] 1/s; eps_r = 12.96; weak CW seed at the
    lowest transition frequency.  dx = 1 mm / 2^20 for every grid size, so
    runs of different length share dt; the end time gives `num_timesteps`
    steps."""
    rng = np.random.default_rng(seed)
    n = 6
    meV = 1e-3 * E0
    spacings = rng.uniform(5.0, 30.0, size=n - 1) * meV * bias
    h_diag = np.concatenate([[0.0], np.cumsum(spacings)])
    tunnel = rng.uniform(0.5, 2.0, size=n - 1) * meV
    h_off, mu_off = [], []
    for col in range(1, n):
        for row in range(col):
            h_off.append(tunnel[row] if col == row + 1 else 0.0)
            if col - row in (1, 2):
                mu_off.append(-E0 * rng.uniform(1.0, 5.0) * 1e-9)
            else:
                mu_off.append(0.0)
    scat = rng.uniform(1e11, 1e13, size=(n, n))
    np.fill_diagonal(scat, 0.0)
    pure = rng.uniform(1e11, 1e13, size=n * (n - 1) // 2)
    qm = {"num_levels": n, "carrier_density": 5e21,
          "h_diag": list(h_diag), "h_off": h_off, "mu_diag": [0.0] * n,
          "mu_off": mu_off, "scattering": scat.tolist(),
          "pure_dephasing": list(pure)}
    eps_r = 12.96
    dx = 1e-3 / ((1 << 20) - 1)
    L = length if length is not None else dx * (num_gridpoints - 1)
    v = C0 / math.sqrt(eps_r)
    dt = 0.5 * dx / v
    # endtime such that ceil(T/dt) + 1 == num_timesteps
    endtime = (num_timesteps - 1.5) * dt
    f_seed = spacings[0] / (2 * PI * HBAR)
    interval = endtime / (num_records - 1) if num_records > 1 else 0.0
    return {
        "name": name, "algo": algo, "num_levels": n,
        "materials": {"AR_" + name: {"qm": qm, "rel_permittivity": eps_r}},
        "regions": [("Active region", "AR_" + name, 0.0, L)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([1.0] + [0.0] * (n - 1), []),
        "sources": [{"kind": "cw", "name": "seed", "position": 0.0,
                     "hard": False, "ampl": cw_ampl, "freq": f_seed,
                     "phase": 0.0}],
        "records": [{"name": "inv12", "interval": interval},
                    {"name": "e", "interval": interval}],
    }


def sit_2lvl_sweep(num_runs=512, seed=1, min_log2=12, max_log2=17,
                   endtime=1.5e-12, num_records=16):
    """Batched two-level sweep: grid lengths log-uniform in
    2^min_log2..2^max_log2 cells with dx fixed to the single-run value
    (150 um / 8191), pulse areas U[0.5 pi, 4 pi], T2 log-uniform in
    10 fs..1 ns; otherwise as sit_2lvl."""
    rng = np.random.default_rng(seed)
    dx = 150e-6 / 8191
    specs = []
    for r in range(num_runs):
        nx = int(round(2 ** rng.uniform(min_log2, max_log2)))
        area = rng.uniform(0.5 * PI, 4 * PI)
        t2 = 10 ** rng.uniform(math.log10(10e-15), math.log10(1e-9))
        L = dx * (nx - 1)
        specs.append(sit_2lvl(num_gridpoints=nx, endtime=endtime, area=area,
                              t2=t2, num_records=num_records, length=L,
                              name="s%03d" % r))
    return specs


def qcl6_sweep(num_runs=64, seed=1, min_log2=14, max_log2=18,
               num_timesteps=4000000, num_records=16):
    """Batched six-level sweep: each run its own seeded Hamiltonian and
    rates, grid lengths log-uniform in 2^14..2^18 cells (shared dx), bias
    scanned by scaling the level spacings over [0.9, 1.1]."""
    rng = np.random.default_rng(seed)
    specs = []
    for r in range(num_runs):
        nx = int(round(2 ** rng.uniform(min_log2, max_log2)))
        bias = 0.9 + 0.2 * r / max(1, num_runs - 1)
        specs.append(qcl6(num_gridpoints=nx, num_timesteps=num_timesteps,
                          seed=seed * 1000 + r, bias=bias,
                          num_records=num_records, name="q%02d" % r))
    return specs


def maxwell_line(num_gridpoints=1 << 27, num_timesteps=4000,
                 name="maxwell"):
    """Pure Maxwell stencil (noop density): a dielectric line (eps_r 4,
    small losses) of num_gridpoints cells with dx = 1 nm, a soft Gaussian
    source at cell 100 and one point record of E; end time chosen for
    num_timesteps steps."""
    dx = 1e-9
    L = dx * (num_gridpoints - 1)
    v = C0 / 2.0
    dt = 0.5 * dx / v
    endtime = (num_timesteps - 1.5) * dt
    return {
        "name": name, "algo": "noop", "num_levels": 0,
        "materials": {"Die": {"rel_permittivity": 4.0, "losses": 10.0}},
        "regions": [("line", "Die", 0.0, L)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([1.0, 0.0], []),
        "sources": [{"kind": "gauss", "name": "g", "position": 100 * dx,
                     "hard": False, "ampl": 1e3, "freq": 1e15,
                     "phase": 500 * dt, "tau": 100 * dt}],
        "records": [{"name": "e", "interval": 0.0,
                     "position": 400 * dx}],
    }


# ---------------------------------------------- reference tool setups -----
def ziolkowski1995(num_gridpoints=32768, endtime=200e-15):
    """mbsolve-tool.cpp:249-309 (self-induced transparency, digitized in
    tools/reference-data/ziolkowski1995-2.csv)."""
    qm = {"type": "2lvl", "carrier_density": 1e24,
          "transition_freq": 2 * PI * 2e14, "dipole_moment": 6.24e-11,
          "scattering_rate": 1.0e10, "dephasing_rate": 1.0e10,
          "equilibrium_inversion": -1.0}
    return {
        "name": "Ziolkowski", "algo": "reg-cayley", "num_levels": 2,
        "materials": {"Vacuum": {}, "AR_Ziolkowski": {"qm": qm}},
        "regions": [("Vacuum left", "Vacuum", 0.0, 7.5e-6),
                    ("Active region", "AR_Ziolkowski", 7.5e-6, 142.5e-6),
                    ("Vacuum right", "Vacuum", 142.5e-6, 150e-6)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([1.0, 0.0], []),
        "sources": [{"kind": "sech", "name": "sech", "position": 0.0,
                     "hard": True, "ampl": 4.2186e9, "freq": 2e14,
                     "phase": 10.0, "beta": 2e14}],
        "records": [{"name": "inv12", "interval": 2.5e-15},
                    {"name": "e", "interval": 2.5e-15}],
    }


def song2005(endtime=80e-15, num_timesteps=10000):
    """mbsolve-tool.cpp:163-248 (V-type three-level system, single grid
    point; digitized in tools/reference-data/song2005-3a.csv)."""
    rate = 1e10
    qm = {"num_levels": 3, "carrier_density": 6e24,
          "h_diag": [0.0, 2.3717 * HBAR * 1e15, 2.4165 * HBAR * 1e15],
          "h_off": [0.0, 0.0, 0.0], "mu_diag": [0.0, 0.0, 0.0],
          "mu_off": [-E0 * 9.2374e-11, -E0 * 9.2374e-11 * math.sqrt(2), 0.0],
          "scattering": [[0, rate, rate], [rate, 0, rate], [rate, rate, 0]],
          "pure_dephasing": [0.0, 0.0, 0.0]}
    return {
        "name": "Song", "algo": "cvr-rodr", "num_levels": 3,
        "materials": {"Vacuum": {}, "AR_Song": {"qm": qm}},
        "regions": [("Active region (single point)", "AR_Song", 0.0, 0.0)],
        "num_gridpoints": 1, "endtime": endtime,
        "num_timesteps": num_timesteps,
        "rho_init": ([1.0, 0.0, 0.0], []),
        "sources": [{"kind": "sech", "name": "sech", "position": 0.0,
                     "hard": True, "ampl": 3.5471e9, "freq": 3.8118e14,
                     "phase": 17.248, "beta": 1.76 / 5e-15,
                     "phase_sin": -PI / 2}],
        "records": [{"name": "d11", "type": 3, "row": 1, "col": 1,
                     "interval": 0.0, "position": 0.0},
                    {"name": "d22", "type": 3, "row": 2, "col": 2,
                     "interval": 0.0, "position": 0.0},
                    {"name": "d33", "type": 3, "row": 3, "col": 3,
                     "interval": 0.0, "position": 0.0},
                    {"name": "e", "interval": 0.0, "position": 0.0}],
    }


def marskar2011(num_levels=6, num_gridpoints=8192, endtime=2e-12):
    """mbsolve-tool.cpp:479-662 (anharmonic ladder, generalised to N
    levels; digitized rho55/rho66 in tools/reference-data)."""
    N = num_levels
    energies = [0.0] * N
    for i in range(1, N):
        energies[i] = energies[i - 1] + (1.0 - 0.1 * (i - 3)) * HBAR * 2 * \
            PI * 1e13
    dipoles = []
    for i in range(1, N):
        dipoles.extend([0.0] * (i - 1))
        dipoles.append(-1e-29)
    T = 600.0
    scat, pure = [], []
    for i in range(N):
        row = []
        for j in range(N):
            if i == j:
                rate = 0.0
            elif i == j - 1:
                rate = 1e12 / (1 + 0.1 * i)
            elif i == j + 1:
                rate = 1e12 / (1 + 0.1 * j) * math.exp(
                    (energies[j] - energies[i]) / (KB * T))
            else:
                rate = 0.0
            row.append(rate)
            pure.append(1e12)
        scat.append(row)
    total = 0.0
    for i in range(N):
        prod = 1.0
        for j in range(i):
            prod *= scat[j + 1][j] / scat[j][j + 1]
        total += prod
    pops = [1.0 / total]
    for i in range(1, N):
        pops.append(pops[i - 1] * scat[i][i - 1] / scat[i - 1][i])
    qm = {"num_levels": N, "carrier_density": 1e25, "h_diag": energies,
          "h_off": [0.0] * (N * (N - 1) // 2), "mu_diag": [0.0] * N,
          "mu_off": dipoles, "scattering": scat,
          "pure_dephasing": pure}
    tau = 100e-15
    return {
        "name": "marskar2011-%dlvl" % N, "algo": "cvr-rodr",
        "num_levels": N,
        "materials": {"Vacuum": {}, "AR_Marsk": {"qm": qm}},
        "regions": [("Active region", "AR_Marsk", 0.0, 1e-3)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": (pops, []),
        "sources": [{"kind": "gauss", "name": "gaussian", "position": 0.0,
                     "hard": True, "ampl": 5e8, "freq": 1e13,
                     "phase": 3.0 * tau, "tau": tau}],
        "records": ([{"name": "d%d%d" % (i, i), "type": 3, "row": i,
                      "col": i, "interval": 0.0, "position": 0.0}
                     for i in range(1, N + 1)] +
                    [{"name": "e", "type": 0, "row": 0, "col": 0,
                      "interval": 0.0, "position": 0.0}]),
    }


def tzenov2016(num_gridpoints=8192, endtime=0.2e-9, seed=1):
    """mbsolve-tool.cpp:310-392 (five-level quantum cascade laser frequency
    comb, 5 mm active region, semi-transparent mirrors 0.8 / 0.8 that no
    solver evaluates, no source: the lasing starts from the random initial
    field, here seeded; E at x = 0 and the five populations at x = 0.1 mm
    recorded every step)."""
    e = E0
    qm = {"num_levels": 5, "carrier_density": 5.6e21,
          "h_diag": [0.10103 * e, 0.09677 * e, 0.09720 * e, 0.08129 * e,
                     0.07633 * e],
          "h_off": [0.0, 1.2329e-3 * e, -1.3447e-3 * e] + [0.0] * 7,
          "mu_diag": [0.0] * 5,
          "mu_off": [0.0] * 5 + [-e * 4e-9] + [0.0] * 4,
          "scattering": [
              [0.0, 0.4947e12, 0.0974e12, 0.8116e12, 1.0410e12],
              [0.8245e12, 0.0, 0.1358e12, 0.6621e12, 1.1240e12],
              [0.0229e12, 0.0469e12, 0.0, 0.0794e12, 0.0357e12],
              [0.0047e12, 0.0029e12, 0.1252e12, 0.0, 0.2810e12],
              [0.0049e12, 0.0049e12, 0.1101e12, 0.4949e12, 0.0]],
          "pure_dephasing": [0.0, 1.0 / 1.6e-12, 1.0 / 1.6e-12] +
                            [1.0 / 4.8e-12] * 7}
    return {
        "name": "tzenov2016", "algo": "cvr-rodr", "num_levels": 5,
        "materials": {"AR": {"qm": qm, "rel_permittivity": 19.35,
                             "overlap_factor": 0.9, "losses": 1300.0}},
        "regions": [("Active region", "AR", 0.0, 5e-3)],
        "reflectivity": (0.8, 0.8),
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([0.0, 0.0, 1.0, 0.0, 0.0], []),
        "e_init_random": seed,
        "sources": [],
        "records": ([{"name": "e0", "type": 0, "row": 1, "col": 1,
                      "interval": 0.0, "position": 0.0}] +
                    [{"name": "d%d%d" % (i, i), "type": 3, "row": i,
                      "col": i, "interval": 0.0, "position": 0.1e-3}
                     for i in range(1, 6)]),
    }


def tzenov2018_cpml(num_gridpoints=8192, endtime=2e-9, algo="reg-cayley"):
    """mbsolve-tool.cpp:394-477 (colliding-pulse mode locking: gain /
    absorber / gain two-level sections; the reference uses a random initial
    field, here a constant zero field plus the initial coherence)."""
    gain = {"type": "2lvl", "carrier_density": 5e21,
            "transition_freq": 2 * PI * 3.4e12, "dipole_moment": 2e-9,
            "scattering_rate": 1.0 / 10e-12, "dephasing_rate": 1.0 / 200e-15,
            "equilibrium_inversion": 1.0}
    absorber = {"type": "2lvl", "carrier_density": 1e21,
                "transition_freq": 2 * PI * 3.4e12, "dipole_moment": 6e-9,
                "scattering_rate": 1.0 / 3e-12,
                "dephasing_rate": 1.0 / 160e-15,
                "equilibrium_inversion": -1.0}
    return {
        "name": "tzenov-cpml", "algo": algo, "num_levels": 2,
        "materials": {"Absorber": {"qm": absorber, "rel_permittivity": 12.96,
                                   "losses": 500.0},
                      "Gain": {"qm": gain, "rel_permittivity": 12.96,
                               "losses": 500.0}},
        "regions": [("Gain R", "Gain", 0.0, 0.5e-3),
                    ("Absorber", "Absorber", 0.5e-3, 0.625e-3),
                    ("Gain L", "Gain", 0.625e-3, 1.125e-3)],
        "num_gridpoints": num_gridpoints, "endtime": endtime,
        "rho_init": ([0.5, 0.5], [0.001]),
        "records": [{"name": "inv12", "interval": 1e-12},
                    {"name": "e", "interval": 1e-12},
                    {"name": "e0", "type": 0, "row": 1, "col": 1,
                     "interval": 0.0, "position": 0.0},
                    {"name": "h0", "type": 2, "row": 1, "col": 1,
                     "interval": 0.0, "position": 1.373e-7}],
    }
