"""flowhip_solve_hydro_2d_static, the step that allocates nothing, against
solve_hydro_2d (the march schedule) on the same inputs: every array, the dt
pair and the totals bit for bit, over two meshes in a row, the second
smaller than the first (the scratch keeps the first call's contents), and a
mesh beyond the scratch refused before any launch."""
import ctypes

import numpy as np
import pytest

import flow_amd
from flow_amd import native, problems

pytestmark = pytest.mark.gpu


def _pair(nx, ny, init, dt):
    out = []
    for _ in range(2):
        m = flow_amd.Mesh(nx, ny, dt=dt)
        st = flow_amd.DeviceState(m)
        st.upload("rho", init[0])
        st.upload("e", init[1])
        st.reflect_initial()
        out.append((m, st))
    return out


@pytest.mark.parametrize("nx,ny,steps", [(150, 97, 5), (61, 38, 4)])
def test_static_step_equals_solve_hydro_2d(nx, ny, steps):
    print(f"case no-alloc step {nx}x{ny} steps={steps}", flush=True)
    rho, e = problems.boxes(nx, ny, [(1.0, 0.25, 0.0, 0.0, 1.0, 1.0),
                                     (1.0, 25.0, 0.3, 0.4, 0.2, 0.15),
                                     (0.5, 5.0, 0.6, 0.1, 0.1, 0.3)])
    (ma, sa), (mb, sb) = _pair(nx, ny, (rho, e), 1.0e-4)
    # the same pre-loop dt for both, from set_timestep on the first
    sa.set_timestep(0)
    mb.set_dt(ma.dt, ma.dt_h)
    for tt in range(steps):
        # solve_hydro_2d + print_conservation, one step
        _, _, mass_a, energy_a, _ = sa.run(1, tt0=tt, pre_timestep=False)
        rc, mass_b, energy_b = sb.solve_static(tt)
        assert rc == 0
        assert (mb.dt, mb.dt_h) == (ma.dt, ma.dt_h), tt
        assert (mass_b, energy_b) == (mass_a[0], energy_a[0]), tt
    for k in ("rho", "e", "u", "v", "mom_x", "mom_y", "F_x", "F_y", "rho_old"):
        a, b = sa.download(k), sb.download(k)
        ndiff = int(np.count_nonzero(a != b))
        print(f"  {k:7s} bitwise_diffs={ndiff}", flush=True)
        assert ndiff == 0, k
    for m, s in ((ma, sa), (mb, sb)):
        s.close()
        m.close()


def test_static_step_refuses_a_mesh_beyond_its_scratch():
    """The size check runs before any launch: the arrays are never read."""
    big = 46336
    s = native.MeshStruct()
    s.local_nx = s.local_ny = big + 4
    s.global_nx = s.global_ny = big
    s.pad = 2
    s.dt = s.dt_h = 1.0e-6
    dummy = ctypes.c_void_p(16)
    rc = native.lib().flowhip_solve_hydro_2d_static(
        ctypes.byref(s), 0, *([dummy] * 16), None, None, None)
    assert rc == -2
