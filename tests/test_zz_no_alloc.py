"""mbh_fetch_state_noalloc against mbh_fetch_state, through the C ABI.

The allocation-free read-out stages every conversion in one fixed device
array per GPU; it must return exactly what mbh_fetch_state returns, for
every state layout (reference FP64 / FP32, Bloch form FP32 / FP64, z-form
FP32 with the equilibrium shift / FP64, no density), on two calls in a row
of which the second reads a smaller run, and across contexts sharing the
array.  Runs too large for the array are rejected with a clear error."""
import numpy as np
import pytest

import mbsolve_amd.lib as mb
from mbsolve_amd import capi, setups

pytestmark = pytest.mark.gpu

DX = 150e-6 / 8191


def _problems(kind, sizes, steps):
    out = []
    for k, n in enumerate(sizes):
        L = (n - 1) * DX
        if kind == "sit2":
            spec = setups.sit_2lvl(num_gridpoints=n, endtime=0.2e-12,
                                   num_records=4, length=L, name="a%d" % k)
        elif kind == "v3":
            spec = setups.v3_rodr(num_gridpoints=n, endtime=0.2e-12,
                                  num_records=4, length=L, seed=k + 1,
                                  name="b%d" % k)
        elif kind == "cayley3":
            spec = setups.v3_rodr(num_gridpoints=n, endtime=0.2e-12,
                                  num_records=4, length=L, seed=k + 1,
                                  name="c%d" % k)
        else:
            spec = setups.maxwell_line(num_gridpoints=n,
                                       num_timesteps=steps, name="m%d" % k)
            spec["records"] = [{"name": "e",
                                "interval": spec["endtime"] / 4}]
        out.append(spec)
    return out


CASES = [
    ("sit2", "hip-fdtd-2lvl-reg-cayley", "fp64"),
    ("sit2", "hip-fdtd-2lvl-reg-cayley-mixed", "mixed"),
    ("sit2", "hip-fdtd-2lvl-reg-cayley-zform", "zform64"),
    ("cayley3", "hip-fdtd-3lvl-reg-cayley-mixed", "mixed"),
    ("v3", "hip-fdtd-3lvl-rodr", "fp64"),
    ("v3", "hip-fdtd-3lvl-rodr-mixed", "mixed"),
    ("v3", "hip-fdtd-3lvl-rodr-mfma64", "mfma64"),
    ("noop", "hip-fdtd-noop-mixed", "mixed"),
]


def _build(specs, name):
    mb.material.clear_library()
    probs = []
    for sp in specs:
        dev, scen = setups.build(sp, clear_library=False)
        probs.append(mb.build_problem(name, dev, scen))
    return probs


def _same(a, b, what):
    for k in ("e", "h", "p", "d"):
        assert a[k].shape == b[k].shape, (what, k)
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.mark.parametrize("kind,name,form", CASES)
def test_noalloc_matches_fetch_state(gpu, kind, name, form):
    specs = _problems(kind, (3000, 1100, 700), steps=300)
    probs = _build(specs, name)
    print("case noalloc", name, form, flush=True)
    # a batch of the two larger runs, read-outs of run 0 then the smaller
    # run 1; then a second context with the smallest run
    ctx = capi.Context(probs[:2], capi.options(form))
    nt = int(probs[0]["num_timesteps"])
    ctx.run(0, nt // 2 + 3)
    for run in (0, 1):
        want = ctx.state(run)
        got = ctx.state(run, noalloc=True)
        assert np.max(np.abs(want["e"])) > 0, (name, run)
        _same(got, want, (name, run))
    ctx.run(nt // 2 + 3, nt)
    got1 = ctx.state(1, noalloc=True)
    _same(got1, ctx.state(1), (name, "end"))
    ctx.close()
    small = capi.Context(probs[2:], capi.options(form))
    small.run(0, nt)
    _same(small.state(0, noalloc=True), small.state(0), (name, "small"))
    small.close()


def test_noalloc_rejects_runs_beyond_the_array(gpu):
    """A run whose H array (N_x + 1 doubles) does not fit the staging array
    is rejected with MBH_ERR_INVALID; mbh_fetch_state still serves it."""
    n = capi.MBH_NOALLOC_SCRATCH_BYTES // 8  # N_x + 1 doubles do not fit
    specs = _problems("noop", (n,), steps=20)
    probs = _build(specs, "hip-fdtd-noop-mixed")
    ctx = capi.Context(probs, capi.options("mixed"))
    ctx.run(0, 20)
    with pytest.raises(capi.MBHError) as ei:
        ctx.state(0, noalloc=True)
    assert ei.value.code == 1
    assert "staging" in str(ei.value)
    st = ctx.state(0)
    assert st["e"].shape == (n,)
    ctx.close()
