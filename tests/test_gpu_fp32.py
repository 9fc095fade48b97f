"""The FP32-state precision form against the oracle (FP64 omp3): per-field
relative L-infinity and L1 error, the dt sequence and the drift of the mass
and energy totals. The tolerances are the precision form's stated bounds;
fields are compared over the regions omp3 defines."""
import numpy as np
import pytest

from flow_amd import problems
from tests import harness

# (problem, n, steps, dt param)
CASES = [
    ("multi_blast", 64, 20, 1.0e-5),
    ("square_blast", 96, 20, 1.0e-5),
    ("sod_x", 128, 40, 1.0e-4),
    ("sod_y", 64, 40, 1.0e-4),
]
FIELD_REL_LINF = 1.0e-4
DT_REL = 1.0e-5
TOTAL_REL = 1.0e-6


def gpu_run_f32(problem_fn, n, steps, dt, init=None, max_dt=1.0e30):
    import flow_amd
    m = flow_amd.Mesh(n, n, dt=dt, max_dt=max_dt)
    st = flow_amd.DeviceState32(m)
    rho, e = init if init is not None else problem_fn(n)
    st.upload("rho", rho)
    st.upload("e", e)
    st.reflect_initial()
    steps_done, dts, mass, energy, _ = st.run(steps)
    out = dict(arrays=st.download_all(), dt=dts, mass=mass, energy=energy,
               steps=steps_done, L=m.nx)
    st.close()
    m.close()
    return out


def field_errors(got, ref, L):
    g, r = harness.views(got, L), harness.views(ref, L)
    reg = harness.state_regions(L)
    out = {}
    for k in ("rho", "e", "u", "v"):
        a, b = g[k][reg[k]], r[k][reg[k]]
        den1 = np.sum(np.abs(b))
        out[k] = (harness.rel_linf(a, b),
                  float(np.sum(np.abs(a - b)) / den1) if den1 else float(np.sum(np.abs(a - b))))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,steps,dt", CASES)
def test_fp32_state_within_stated_error(name, n, steps, dt):
    fn = problems.PROBLEMS[name]
    print(f"case fp32 {name} n={n} steps={steps} dt={dt}", flush=True)
    ref = harness.oracle_run(fn, n, steps, dt=dt)
    got = gpu_run_f32(fn, n, steps, dt)
    assert got["steps"] == ref["steps"] == steps
    errs = field_errors(got["arrays"], ref["arrays"], ref["L"])
    for k, (linf, l1) in errs.items():
        print(f"  {k:4s} rel_linf={linf:.3e} rel_l1={l1:.3e}", flush=True)
    dt_rel = np.max(np.abs(got["dt"] - ref["dt"]) / np.abs(ref["dt"]))
    m_rel = np.max(np.abs(got["mass"] - ref["mass"]) / np.abs(ref["mass"]))
    e_rel = np.max(np.abs(got["energy"] - ref["energy"]) / np.abs(ref["energy"]))
    print(f"  dt_rel={dt_rel:.3e} mass_rel={m_rel:.3e} energy_rel={e_rel:.3e}", flush=True)
    for k, (linf, _) in errs.items():
        assert linf <= FIELD_REL_LINF, (k, linf)
    assert dt_rel <= DT_REL
    assert m_rel <= TOTAL_REL and e_rel <= TOTAL_REL


@pytest.mark.gpu
def test_fp32_uniform_state_stays_at_rest():
    """A uniform state at rest: every flux is exactly zero, so rho is
    unchanged bit for bit and u = v = 0 exactly; e = (rho_old * e) / rho
    may round by one FP32 unit per step."""
    n = 40
    L = n + 4
    rho = np.full((L + 1) * (L + 1), 0.7)
    e = np.full((L + 1) * (L + 1), 3.25)
    steps = 6
    got = gpu_run_f32(None, n, steps, 1.0e-4, init=(rho, e))
    v = harness.views(got["arrays"], L)
    inner = (slice(2, L - 2), slice(2, L - 2))
    assert np.all(v["rho"][inner] == np.float32(0.7))
    rel = np.abs(v["e"][inner] - np.float32(3.25)) / 3.25
    assert rel.max() <= steps * 2.0 ** -23
    assert np.all(v["u"] == 0.0) and np.all(v["v"] == 0.0)


@pytest.mark.gpu
def test_fp32_near_vacuum_cells_stay_finite():
    """A blast next to a near-vacuum region (rho = 1e-8 < 1e-6: the EOS of
    those cells runs in FP64)."""
    n = 48
    rho, e = problems.square_blast(n)
    L = n + 4
    r2 = rho.copy()
    r2[: L * L].reshape(L, L)[:, : L // 3] = 1.0e-8
    init = (r2, e)
    ref = harness.oracle_run(None, n, 10, dt=1.0e-5, init=init)
    got = gpu_run_f32(None, n, 10, 1.0e-5, init=init)
    for k, (linf, l1) in field_errors(got["arrays"], ref["arrays"], ref["L"]).items():
        print(f"  {k:4s} rel_linf={linf:.3e} rel_l1={l1:.3e}", flush=True)
        assert linf <= FIELD_REL_LINF, (k, linf)
    assert np.all(np.isfinite(got["arrays"]["e"]))


@pytest.mark.gpu
def test_fp32_state_binding_max_dt():
    """max_dt below the CFL step: the FP32 form compares squared candidates
    starting from max_dt^2 and takes one square root of the minimum, so its
    dt sequence still follows omp3's (omp3/flow.c:80, 103-105)."""
    n, steps, dt, max_dt = 64, 10, 1e-4, 2e-5
    print(f"case fp32 max_dt multi_blast n={n} steps={steps} max_dt={max_dt}", flush=True)
    ref = harness.oracle_run(problems.multi_blast, n, steps, max_dt=max_dt, dt=dt)
    got = gpu_run_f32(problems.multi_blast, n, steps, dt, max_dt=max_dt)
    dt_rel = np.max(np.abs(got["dt"] - ref["dt"]) / np.abs(ref["dt"]))
    print(f"  dt_rel={dt_rel:.3e}", flush=True)
    assert dt_rel <= DT_REL
    errs = field_errors(got["arrays"], ref["arrays"], ref["L"])
    for k, (linf, _) in errs.items():
        assert linf <= FIELD_REL_LINF, (k, linf)
