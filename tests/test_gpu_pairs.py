"""The Lagrangian and mass/energy marches with two columns per lane give the
fields of the one-column marches bit for bit, FP64 and FP32 state.

On uniform power-of-two meshes with an even padded width the march
schedule runs the interior items of those two marches two columns per lane
(strips of 120 and 124 columns); edge items go through the one-column form
in sub-strips. By default it does so from 2048^2 interior cells up;
FLOWHIP_PAIRS=1 (read when a mesh is created) asks for it on every mesh,
which this test uses so that small and rectangular meshes take the form
too, and FLOWHIP_PAIRS=0 keeps one column per lane. Each case runs both
ways over both sweep orders; every
array and the dt sequence must agree bit for bit (the CFL minimum does not
depend on the order the candidates are taken in), and the mass and energy
totals to 1e-12: each lane's sums then take two columns a row over other
chunk heights, so the rounding of the per-item partial sums differs.
"""
import numpy as np
import pytest

import flow_amd
from flow_amd import problems
from tests import harness

pytestmark = pytest.mark.gpu

# (sides, steps, dt param), all power-of-two sides (so the pair forms run):
# every mesh's last strip is short of 124 (120) columns and goes through the
# one-column form; 256 x 512 and 2048 x 1024 are rectangular.
CASES = [((1024, 1024), 6, 1.0e-5), ((256, 512), 6, 1.0e-5), ((2048, 1024), 4, 1.0e-5),
         ((8192, 8192), 2, 1.0e-6)]


def _init(n):
    nx, ny = n
    if nx == ny:
        return problems.multi_blast(nx)
    return problems.boxes(nx, ny, [(1.0, 0.25, 0.0, 0.0, 1.0, 1.0),
                                   (1.0, 250.0, 0.45, 0.45, 0.1, 0.1),
                                   (0.5, 50.0, 0.1, 0.6, 0.2, 0.15)])


def _run(monkeypatch, pairs, n, steps, dt, fp32):
    monkeypatch.setenv("FLOWHIP_PAIRS", "1" if pairs else "0")
    rho, e = _init(n)
    if not fp32:
        return harness.gpu_run(None, n, steps, flow_amd.MARCH, dt=dt, init=(rho, e))
    m = flow_amd.Mesh(n[0], n[1], dt=dt)
    st = flow_amd.DeviceState32(m)
    st.upload("rho", rho)
    st.upload("e", e)
    st.reflect_initial()
    steps_done, dts, mass, energy, _ = st.run(steps)
    out = dict(arrays=st.download_all(), dt=dts, mass=mass, energy=energy, steps=steps_done)
    st.close()
    m.close()
    return out


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("n,steps,dt", CASES)
def test_pairs_equal_one_column(monkeypatch, n, steps, dt, fp32):
    print(f"case pairs n={n} steps={steps} fp32={fp32}", flush=True)
    one = _run(monkeypatch, False, n, steps, dt, fp32)
    two = _run(monkeypatch, True, n, steps, dt, fp32)
    assert one["steps"] == two["steps"] == steps
    for k, a in one["arrays"].items():
        ndiff = int(np.count_nonzero(two["arrays"][k].view(np.uint8) != a.view(np.uint8)))
        assert ndiff == 0, (k, ndiff)
    np.testing.assert_array_equal(two["dt"], one["dt"])
    np.testing.assert_allclose(two["mass"], one["mass"], rtol=1e-12)
    np.testing.assert_allclose(two["energy"], one["energy"], rtol=1e-12)
