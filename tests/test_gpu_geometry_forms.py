"""The three geometry forms of the march kernels give the same bits.

On a mesh whose spacings are powers of two the march kernels replace
quotients by spacings with products by exact reciprocals, and when the
spacings are also uniform the mass/energy and momentum marches read the row
geometry once per work item. FLOWHIP_UNIFORM=0 and FLOWHIP_POW2=0 (read
when a mesh is created) switch those forms off. Each problem here runs in
all three forms, FP64 and FP32 state, and every array and the dt sequence
must agree bit for bit; the FP64 default form is also checked against the
oracle. The totals must agree to 1e-12: on large uniform power-of-two
meshes the Lagrangian and mass/energy marches hold two columns per lane,
whose per-item partial sums round differently from the one-column form's.
"""
import numpy as np
import pytest

import flow_amd
from flow_amd import problems
from tests import harness

pytestmark = pytest.mark.gpu

FORMS = {"uniform": {}, "pow2": {"FLOWHIP_UNIFORM": "0"}, "ieee": {"FLOWHIP_POW2": "0"}}
# (problem, sides, steps, dt param); 256^2 and 128 x 256 are powers of two,
# 192 x 96 is not (every form is then the IEEE one)
CASES = [("multi_blast", (256, 256), 6, 1.0e-5), ("sod_y", (128, 256), 8, 1.0e-4),
         ("square_blast", (192, 96), 6, 1.0e-5)]
BOXES = {"sod_y": [(0.125, 2.0, 0.0, 0.0, 1.0, 1.0), (1.0, 2.5, 0.0, 0.0, 1.0, 0.5)],
         "square_blast": [(1.0, 0.25, 0.0, 0.0, 1.0, 1.0),
                          (1.0, 250.0, 0.45, 0.45, 0.1, 0.1)]}


def _init(name, n):
    if name == "multi_blast":
        return problems.multi_blast(n[0])
    return problems.boxes(n[0], n[1], BOXES[name])


def _run(monkeypatch, env, name, n, steps, dt, fp32):
    for k in ("FLOWHIP_UNIFORM", "FLOWHIP_POW2"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    init = _init(name, n)
    if fp32:
        nx, ny = n
        m = flow_amd.Mesh(nx, ny, dt=dt)
        st = flow_amd.DeviceState32(m)
        st.upload("rho", init[0])
        st.upload("e", init[1])
        st.reflect_initial()
        steps_done, dts, mass, energy, _ = st.run(steps)
        out = dict(arrays=st.download_all(), dt=dts, mass=mass, energy=energy,
                   steps=steps_done)
        st.close()
        m.close()
        return out
    return harness.gpu_run(None, n, steps, flow_amd.MARCH, dt=dt, init=init)


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("name,n,steps,dt", CASES)
def test_geometry_forms_agree(monkeypatch, name, n, steps, dt, fp32):
    print(f"case {name} n={n} steps={steps} fp32={fp32}", flush=True)
    runs = {form: _run(monkeypatch, env, name, n, steps, dt, fp32)
            for form, env in FORMS.items()}
    base = runs["uniform"]
    assert base["steps"] == steps
    for form in ("pow2", "ieee"):
        got = runs[form]
        for k, a in base["arrays"].items():
            ndiff = int(np.count_nonzero(got["arrays"][k].view(np.uint8) !=
                                         a.view(np.uint8)))
            assert ndiff == 0, (form, k, ndiff)
        np.testing.assert_array_equal(got["dt"], base["dt"])
        np.testing.assert_allclose(got["mass"], base["mass"], rtol=1e-12)
        np.testing.assert_allclose(got["energy"], base["energy"], rtol=1e-12)
    if not fp32:
        ref = harness.oracle_run(None, n, steps, dt=dt, init=_init(name, n))
        for k, (rel, ndiff) in harness.compare_state(base["arrays"], ref["arrays"],
                                                     ref["L"]).items():
            assert ndiff == 0, (k, rel, ndiff)
        np.testing.assert_array_equal(base["dt"], ref["dt"])
