"""set_timestep's two rules for undefined CFL candidates, on the GPU,
against the oracle under the same rule.

- The pre-loop set_timestep (k_timestep / k_timestep_serial) on crafted and
  random states with NaN candidates: dt and dt_h bit for bit, and the count
  of undefined candidates equal to the oracle's.
- Whole runs under omp3's serial rule on every schedule: a 96^2 copy of the
  4096^2 blast with its .params dt scaled so that the first step runs as
  close to the CFL limit as the 4096^2 run does (0.87 of it). Shock heating
  drives cells to negative energy in step 0, the two rules part at the next
  dt, and under the serial rule the state goes non-finite by step 10, as
  omp3's does. Fields are compared with NaN on both sides counted equal.
- The 4096^2 blast itself under the serial rule, first 12 steps.
"""
import numpy as np
import pytest

import flow_amd
from flow_amd import problems
from tests import harness
from tests.test_cfl_rules import make_state

pytestmark = pytest.mark.gpu

RULES = (flow_amd.CFL_DEFINED, flow_amd.CFL_OMP3_SERIAL)


def _states():
    """(name, nx, ny, e, rho, qxx, qyy, max_dt) cases."""
    out = []
    nx, ny = 8, 7
    e = np.linspace(1.0, 2.0, nx * ny).reshape(ny, nx)
    e[1, 3], e[5, 6], e[4, 2] = 9.0, 4.0, -0.5
    out.append(("nan_after_min", nx, ny, e, None, None, None, 1.0e30))
    e = np.full((5, 6), 2.0)
    e[4, 5] = -1.0
    out.append(("nan_last_cell", 6, 5, e, None, None, None, 0.3))
    for seed, (nx, ny, frac) in enumerate([(13, 9, 0.15), (300, 200, 1e-3),
                                           (1500, 1100, 2e-5), (257, 129, 0.5)]):
        rng = np.random.default_rng(500 + seed)
        e = rng.uniform(0.2, 6.0, (ny, nx))
        e[rng.random((ny, nx)) < frac] = -1.0
        rho = rng.uniform(0.5, 2.0, (ny, nx))
        qxx = rng.uniform(0.0, 0.5, (ny, nx))
        qyy = rng.uniform(0.0, 0.5, (ny, nx))
        out.append((f"random_{nx}x{ny}", nx, ny, e, rho, qxx, qyy, 1.0e30))
    return out


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("case", _states(), ids=lambda c: c[0])
def test_set_timestep_matches_oracle(oracle, case, rule):
    name, nx, ny, e, rho, qxx, qyy, max_dt = case
    print(f"case set_timestep {name} rule={rule}", flush=True)
    po = oracle
    m, a = make_state(po, nx, ny, e, rho, qxx, qyy, max_dt=max_dt, dt_h=2.0e-3)
    with po.cfl_rule(rule):
        g = po.set_timestep(m, a, 0)
    want = (m.m.dt, m.m.dt_h, po.lib().orc_last_undefined())
    gm = flow_amd.Mesh(nx, ny, max_dt=max_dt)
    gm.set_cfl_rule(rule)
    gm.set_dt(1.0, 2.0e-3)
    st = flow_amd.DeviceState(gm)
    for k in ("rho", "e", "Qxx", "Qyy"):
        st.upload(k, a[k])
    st.set_timestep(0)
    got = (gm.dt, gm.dt_h, gm.undefined_candidates())
    print(f"  local_min={g!r} oracle={want} gpu={got}", flush=True)
    st.close()
    gm.close()
    assert got == want


@pytest.mark.parametrize("schedule,graphs", [(flow_amd.MARCH, False),
                                             (flow_amd.MARCH, True),
                                             (flow_amd.FUSED, False),
                                             (flow_amd.STAGEWISE, False)])
def test_serial_rule_run_matches_oracle(schedule, graphs):
    n, steps, dt = 96, 14, 4.3e-3
    print(f"case serial-rule blast n={n} steps={steps} schedule={schedule} "
          f"graphs={graphs}", flush=True)
    fn = problems.blast_4096
    ref = harness.oracle_run(fn, n, steps, dt=dt, cfl_rule=flow_amd.CFL_OMP3_SERIAL)
    other = harness.oracle_run(fn, n, steps, dt=dt, cfl_rule=flow_amd.CFL_DEFINED)
    # the case discriminates: the rules part at the dt step 0 sets
    assert ref["dt"][0, 0] == other["dt"][0, 0]
    assert ref["dt"][1, 0] != other["dt"][1, 0]
    assert ref["undefined"][1] > 0
    got = harness.gpu_run(fn, n, steps, schedule, dt=dt, graphs=graphs,
                          cfl_rule=flow_amd.CFL_OMP3_SERIAL)
    print(f"  undefined per step: {got['undefined'].tolist()}", flush=True)
    np.testing.assert_array_equal(got["undefined"], ref["undefined"])
    np.testing.assert_array_equal(got["dt"], ref["dt"])
    cmp = harness.compare_state(got["arrays"], ref["arrays"], ref["L"], nan_equal=True)
    for k, (rel, ndiff) in cmp.items():
        print(f"  {k:6s} rel_linf={rel:.3e} diffs={ndiff}", flush=True)
        assert ndiff == 0, (k, ndiff)
    assert np.isnan(ref["arrays"]["rho"]).any()  # omp3's run has diverged


def test_blast_4096_serial_rule_first_12_steps():
    print("case blast_4096 n=4096 steps=12 march serial rule", flush=True)
    fn = problems.blast_4096
    ref = harness.oracle_run(fn, 4096, 12, dt=1e-4, cfl_rule=flow_amd.CFL_OMP3_SERIAL)
    got = harness.gpu_run(fn, 4096, 12, flow_amd.MARCH, dt=1e-4,
                          cfl_rule=flow_amd.CFL_OMP3_SERIAL)
    print(f"  dt: {got['dt'][:, 0].tolist()}", flush=True)
    print(f"  undefined per step: {got['undefined'].tolist()}", flush=True)
    np.testing.assert_array_equal(got["dt"], ref["dt"])
    np.testing.assert_array_equal(got["undefined"], ref["undefined"])
    cmp = harness.compare_state(got["arrays"], ref["arrays"], ref["L"], nan_equal=True)
    for k, (rel, ndiff) in cmp.items():
        print(f"  {k:6s} rel_linf={rel:.3e} diffs={ndiff}", flush=True)
        assert ndiff == 0, (k, ndiff)
    # omp3's dt collapses in step 9, once NaN has spread through the blast
    assert got["dt"][10, 0] < 1e-60
