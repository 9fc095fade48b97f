"""Parity of the hip kernel set with the oracle on identical inputs.

The FP64 bar: rho, e, u and v within relative L-inf
1e-11 after N steps, the dt sequence within 1e-12 relative, the mass and
energy totals to round-off. The kernels keep omp3's expression trees with
contraction off, so the tests also count bitwise differences, which are
expected to be zero.
"""
import numpy as np
import pytest

import flow_amd
from flow_amd import problems
from tests import harness

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-11
DT_TOL = 1e-12
SUM_TOL = 1e-13

# (problem, n, steps, .params dt). The last case reproduces what flow.params
# does at 4096^2: its first step drives e negative in some cells, whose CFL
# candidates are NaN and must be left out of the dt minimum.
CASES = [
    ("sod_x", 32, 3, 1e-4), ("sod_y", 32, 3, 1e-4),
    ("square_blast", 40, 4, 1e-4), ("multi_blast", 64, 5, 1e-5),
    ("multi_blast", 100, 2, 1e-5), ("flow_params", 50, 6, 1e-4),
    ("flow_params", 64, 4, 0.064),
]


def _check(got, ref, L, bitwise):
    cmp = harness.compare_state(got["arrays"], ref["arrays"], L)
    for k, (rel, ndiff) in cmp.items():
        print(f"  {k:6s} rel_linf={rel:.3e} bitwise_diffs={ndiff}", flush=True)
        assert rel <= FIELD_TOL, (k, rel)
        if bitwise:
            assert ndiff == 0, (k, ndiff)
    assert got["steps"] == ref["steps"]
    dt_rel = np.max(np.abs(got["dt"] - ref["dt"]) / np.abs(ref["dt"]))
    assert dt_rel <= DT_TOL, dt_rel
    if bitwise:
        np.testing.assert_array_equal(got["dt"], ref["dt"])
    np.testing.assert_allclose(got["mass"], ref["mass"], rtol=SUM_TOL)
    np.testing.assert_allclose(got["energy"], ref["energy"], rtol=SUM_TOL)


@pytest.mark.parametrize("name,n,steps,dt", CASES)
def test_stagewise_matches_oracle_bitwise(name, n, steps, dt):
    """The stagewise schedule writes every array omp3 writes: all 16 arrays
    must equal the oracle's bit for bit, ghosts and corners included."""
    print(f"case stagewise {name} n={n} steps={steps} dt={dt}", flush=True)
    fn = problems.PROBLEMS[name]
    ref = harness.oracle_run(fn, n, steps, dt=dt)
    got = harness.gpu_run(fn, n, steps, flow_amd.STAGEWISE, dt=dt)
    _check(got, ref, ref["L"], bitwise=True)
    for k in flow_amd.ARRAY_NAMES:
        diff = int(np.count_nonzero(got["arrays"][k] != ref["arrays"][k]))
        assert diff == 0, (k, diff)


def test_stagewise_kat_flow_params():
    """flow.tests:1 on the GPU: flow.params, 4096^2, 100 iterations."""
    import json, os
    here = os.path.dirname(os.path.abspath(__file__))
    kat = json.load(open(os.path.join(here, "golden", "flow_tests_kat.json")))
    n = kat["nx"]
    print(f"case stagewise KAT n={n}", flush=True)
    got = harness.gpu_run(problems.flow_params_problem, n, kat["iterations"],
                          flow_amd.STAGEWISE, dt=kat["dt"])
    L = got["L"]
    v = harness.views(got["arrays"], L)
    dsum = float(v["rho"][2:L - 2, 2:L - 2].sum())
    esum = float(v["e"][2:L - 2, 2:L - 2].sum())
    print(f"  density {dsum:.12e} energy {esum:.12e}", flush=True)
    assert "%.12e" % dsum == "%.12e" % kat["density"]
    assert "%.12e" % esum == "%.12e" % kat["energy"]


# Fused schedule: the carried state (rho, e, u, v and the momenta on their
# face domains), the mass fluxes and rho_old equal omp3's bit for bit. The
# remaining scratch arrays hold the fused kernels' ping-pong fields.
FUSED_CASES = CASES + [
    ("multi_blast", 2, 3, 1e-5), ("multi_blast", 3, 4, 1e-5),
    ("square_blast", 33, 5, 1e-4), ("sod_x", 47, 4, 1e-4),
    ("multi_blast", 1000, 3, 1e-5),
]


@pytest.mark.parametrize("schedule", ["fused", "march"])
@pytest.mark.parametrize("name,n,steps,dt", FUSED_CASES)
def test_fused_matches_oracle_bitwise(name, n, steps, dt, schedule):
    print(f"case {schedule} {name} n={n} steps={steps} dt={dt}", flush=True)
    fn = problems.PROBLEMS[name]
    ref = harness.oracle_run(fn, n, steps, dt=dt)
    sched = flow_amd.FUSED if schedule == "fused" else flow_amd.MARCH
    got = harness.gpu_run(fn, n, steps, sched, dt=dt, march_rows=16)
    _check(got, ref, ref["L"], bitwise=True)
    L = ref["L"]
    for k in ("F_x", "F_y", "rho_old"):
        a, b = got["arrays"][k], ref["arrays"][k]
        if k == "F_x":
            m = harness.noncorner_mask(L, L + 1).ravel()
            a, b = a[: L * (L + 1)][m], b[: L * (L + 1)][m]
        elif k == "F_y":
            m = harness.noncorner_mask(L + 1, L).ravel()
            a, b = a[: (L + 1) * L][m], b[: (L + 1) * L][m]
        diff = int(np.count_nonzero(a != b))
        print(f"  {k:7s} bitwise_diffs={diff}", flush=True)
        assert diff == 0, (k, diff)


@pytest.mark.parametrize("schedule", ["fused", "march"])
def test_fused_is_deterministic(schedule):
    print(f"case {schedule} determinism n=300", flush=True)
    sched = flow_amd.FUSED if schedule == "fused" else flow_amd.MARCH
    a = harness.gpu_run(problems.multi_blast, 300, 4, sched, dt=1e-5)
    b = harness.gpu_run(problems.multi_blast, 300, 4, sched, dt=1e-5)
    for k in flow_amd.ARRAY_NAMES:
        assert np.array_equal(a["arrays"][k], b["arrays"][k], equal_nan=True), k
    np.testing.assert_array_equal(a["mass"], b["mass"])
    np.testing.assert_array_equal(a["energy"], b["energy"])


@pytest.mark.parametrize("schedule", ["fused", "march"])
def test_fused_kat_flow_params(schedule):
    """flow.tests:1 through the fused schedules."""
    import json, os
    here = os.path.dirname(os.path.abspath(__file__))
    kat = json.load(open(os.path.join(here, "golden", "flow_tests_kat.json")))
    n = kat["nx"]
    print(f"case {schedule} KAT n={n}", flush=True)
    sched = flow_amd.FUSED if schedule == "fused" else flow_amd.MARCH
    got = harness.gpu_run(problems.flow_params_problem, n, kat["iterations"],
                          sched, dt=kat["dt"])
    L = got["L"]
    v = harness.views(got["arrays"], L)
    dsum = float(v["rho"][2:L - 2, 2:L - 2].sum())
    esum = float(v["e"][2:L - 2, 2:L - 2].sum())
    print(f"  density {dsum:.12e} energy {esum:.12e}", flush=True)
    assert "%.12e" % dsum == "%.12e" % kat["density"]
    assert "%.12e" % esum == "%.12e" % kat["energy"]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["stagewise", "fused", "march"])
def test_graph_replay_matches_oracle_bitwise(schedule):
    """Steps replayed from captured hipGraphs (each replay followed by the
    eager dt/totals transfers of the interface) equal the oracle bit for
    bit; a second state on the same mesh forces a recapture."""
    import flow_amd
    sched = {"stagewise": flow_amd.STAGEWISE, "fused": flow_amd.FUSED,
             "march": flow_amd.MARCH}[schedule]
    n, steps, dt = 64, 9, 1.0e-5
    m = flow_amd.Mesh(n, n, dt=dt)
    m.set_schedule(sched)
    m.set_graphs(True)
    for fn in (problems.multi_blast, problems.square_blast):
        print(f"case graphs {schedule} {fn.__name__} n={n} steps={steps}", flush=True)
        ref = harness.oracle_run(fn, n, steps, dt=dt)
        st = flow_amd.DeviceState(m)
        rho, e = fn(n)
        st.upload("rho", rho)
        st.upload("e", e)
        st.reflect_initial()
        m.set_dt(0.3 * dt, 0.3 * dt)
        got_steps, dts, mass, energy, _ = st.run(steps)
        got = st.download_all()
        st.close()
        assert got_steps == steps
        assert np.array_equal(dts, ref["dt"])
        cmp = harness.compare_state(got, ref["arrays"], ref["L"])
        for k, (rel, ndiff) in cmp.items():
            assert ndiff == 0, (k, rel, ndiff)
        np.testing.assert_allclose(mass, ref["mass"], rtol=1e-15, atol=0)
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["stagewise", "fused", "march"])
def test_binding_max_dt_matches_oracle_bitwise(schedule):
    """max_dt below the CFL step (2e-5 on the 64^2 multi-blast, whose CFL
    step grows to ~1.6e-4): the minimum starts from max_dt
    (omp3/flow.c:80) and the dt recurrence settles at C_T * max_dt."""
    sched = {"stagewise": flow_amd.STAGEWISE, "fused": flow_amd.FUSED,
             "march": flow_amd.MARCH}[schedule]
    n, steps, dt, max_dt = 64, 10, 1e-4, 2e-5
    print(f"case max_dt {schedule} multi_blast n={n} steps={steps} max_dt={max_dt}",
          flush=True)
    ref = harness.oracle_run(problems.multi_blast, n, steps, max_dt=max_dt, dt=dt)
    assert ref["dt"][-1][0] == pytest.approx(0.3 * max_dt, rel=1e-12)
    got = harness.gpu_run(problems.multi_blast, n, steps, sched, max_dt=max_dt, dt=dt)
    _check(got, ref, ref["L"], bitwise=True)
