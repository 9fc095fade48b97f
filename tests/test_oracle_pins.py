"""Pins of the oracle (the CPU restatement of omp3) against the reference's
own known answer, textbook closed forms, hand-computed stencils and
invariants of the whole step. Each check is chosen so that a plausible slip
(dropped term, wrong sign, wrong index, swapped operands) fails it.
"""
import json
import math
import os

import numpy as np
import pytest

from flow_amd import problems
from tests.harness import PAD

HERE = os.path.dirname(os.path.abspath(__file__))


def grid_arrays(po, n, width=1.0, dt=1.0e-4):
    m = po.Mesh(n, n, width=width, height=width, dt=dt)
    return m, po.new_arrays(m)


def as2d(a, rows, cols):
    return a[: rows * cols].reshape(rows, cols)


# ---------------------------------------------------------------- the KAT ---

@pytest.mark.slow
def test_flow_tests_known_answer(oracle):
    """flow.tests:1 -- after 100 iterations of flow.params (4096^2, dt 1e-3)
    validate() sums interior density and specific energy. The oracle must
    print the same 13 significant digits (and so pass the reference's 1e-5
    tolerance by a wide margin)."""
    with open(os.path.join(HERE, "golden", "flow_tests_kat.json")) as f:
        kat = json.load(f)
    po = oracle
    n = kat["nx"]
    m = po.Mesh(n, n, width=kat["width"], height=kat["height"],
                max_dt=kat["max_dt"], sim_end=kat["sim_end"], dt=kat["dt"])
    a = po.new_arrays(m)
    for (rho, e, xp, yp, w, h) in kat["problems"]:
        po.rasterise_box(m, a, rho, e, xp, yp, w, h)
    po.reflect_initial(m, a)
    L = m.nx
    steps, _, _, _, _ = po.run(m, a, kat["iterations"])
    assert steps == kat["iterations"]
    rho = as2d(a["rho"], L, L)[PAD:L - PAD, PAD:L - PAD]
    e = as2d(a["e"], L, L)[PAD:L - PAD, PAD:L - PAD]
    dsum, esum = float(rho.sum()), float(e.sum())
    assert "%.12e" % dsum == "%.12e" % kat["density"]
    assert "%.12e" % esum == "%.12e" % kat["energy"]
    assert abs(esum - kat["energy"]) <= 1e-5 * kat["energy"]


def test_kat_mass_pins_rasterisation():
    """The KAT's density equals the initial mass only if problem_1 covers
    819 x 819 cells of the 4096^2 grid -- the left-edge rule; a cell-centre
    rule would give 820^2."""
    with open(os.path.join(HERE, "golden", "flow_tests_kat.json")) as f:
        kat = json.load(f)
    n = kat["nx"]
    ex = problems.edges(n)[PAD:PAD + n]
    x0, w = 0.4, 0.2
    left = int(np.count_nonzero((ex >= x0) & (ex < x0 + w)))
    centre = int(np.count_nonzero((ex + 0.5 / n >= x0) & (ex + 0.5 / n < x0 + w)))
    assert left == 819 and centre == 820
    assert 0.125 * n * n + 0.875 * left * left == kat["density"]


# ---------------------------------------------------------------- minmod ---

@pytest.mark.parametrize("a,b,want", [
    (1.0, 2.0, 1.0), (2.0, 1.0, 1.0), (-1.0, -3.0, -1.0), (-3.0, -0.5, -0.5),
    (1.0, -2.0, 0.0), (-1.0, 2.0, 0.0), (0.0, 5.0, 0.0), (5.0, 0.0, 0.0),
    (2.0, 2.0, 2.0)])
def test_minmod_textbook(oracle, a, b, want):
    assert oracle.lib().orc_minmod(a, b) == want


def test_min_propagates_nan_only_from_second(oracle):
    L = oracle.lib()
    assert L.orc_min(1.0, 2.0) == 1.0 and L.orc_min(2.0, 1.0) == 1.0
    assert math.isnan(L.orc_min(1.0, float("nan")))
    assert L.orc_min(float("nan"), 1.0) == 1.0


# -------------------------------------------------------------- boundary ---

def test_boundary_mirror_maps(oracle):
    """pad-deep mirror strips; corners untouched; INVERT_X negates only the
    east/west copies, INVERT_Y only north/south. Expected values written out
    by index for a 7 x 6 array."""
    nx, ny, pad = 7, 6, 2
    base = np.arange(nx * ny, dtype=np.float64) + 100.0
    for inv, sx, sy in ((0, 1, 1), (1, -1, 1), (2, 1, -1)):
        a = base.copy()
        oracle.boundary(nx, ny, pad, a, inv)
        A, B = a.reshape(ny, nx), base.reshape(ny, nx)
        want = B.copy()
        for j in range(pad, nx - pad):
            want[1, j], want[0, j] = sy * B[2, j], sy * B[3, j]
            want[4, j], want[5, j] = sy * B[3, j], sy * B[2, j]
        for i in range(pad, ny - pad):
            want[i, 1], want[i, 0] = sx * B[i, 2], sx * B[i, 3]
            want[i, 5], want[i, 6] = sx * B[i, 4], sx * B[i, 3]
        np.testing.assert_array_equal(A, want)


# ------------------------------------------------------- Lagrangian stages ---

def test_eos_sod_states(oracle):
    """p = (gamma-1) rho e: the Sod states rho=1, e=2.5 -> p=1 and
    rho=0.125, e=2 -> p=0.1 (to one rounding)."""
    n = 4
    L = n + 2 * PAD
    rho = np.zeros((L + 1) ** 2)
    e = np.zeros_like(rho)
    P = np.zeros_like(rho)
    rho[:L * L] = 1.0
    e[:L * L] = 2.5
    rho[0], e[0] = 0.125, 2.0
    oracle.lib().orc_equation_of_state(L, L, oracle.ptr(P), oracle.ptr(rho),
                                       oracle.ptr(e))
    assert abs(P[1] - 1.0) <= 3e-16 and abs(P[0] - 0.1) <= 3e-16 * 0.1
    assert P[L * L] == 0.0  # nothing written past the cell array


def test_pressure_acceleration_linear_pressure(oracle):
    """Uniform density, pressure linear in x and y: every face momentum
    changes by -dt * dp/dx exactly (to rounding) and u = mom / rho."""
    po = oracle
    n = 8
    m, a = grid_arrays(po, n)
    L = m.nx
    dx = 1.0 / n
    x = (np.arange(L) - PAD + 0.5) * dx
    y = x
    bx, by = 3.0, -2.0
    Pm = 5.0 + bx * x[None, :] + by * y[:, None]
    as2d(a["P"], L, L)[:] = Pm
    as2d(a["rho"], L, L)[:] = 2.0
    dt = 1e-3
    po.lib().orc_pressure_acceleration(L, L, m.ref, dt, po.ptr(a["mom_x"]),
                                       po.ptr(a["mom_y"]), po.ptr(a["u"]),
                                       po.ptr(a["v"]), po.ptr(a["P"]),
                                       po.ptr(a["rho"]))
    mx = as2d(a["mom_x"], L, L + 1)[PAD:L - PAD, PAD:L + 1 - PAD]
    my = as2d(a["mom_y"], L + 1, L)[PAD:L + 1 - PAD, PAD:L - PAD]
    np.testing.assert_allclose(mx, -dt * bx, rtol=1e-12)
    np.testing.assert_allclose(my, -dt * by, rtol=1e-12)
    u = as2d(a["u"], L, L + 1)[PAD:L - PAD, PAD:L + 1 - PAD]
    np.testing.assert_allclose(u, -dt * bx / 2.0, rtol=1e-12)


def test_viscous_stress_hand_stencil(oracle):
    """One compressive stencil computed by hand: faces (0, 0, -1, -3) give
    u_i = -1, u_ii = 0.5(|-2+1| + |-1-0|) = 1, Qxx = -3 rho u_i u_ii = 3 rho.
    Expansion (increasing u) gives zero stress."""
    po = oracle
    n = 4
    m, a = grid_arrays(po, n)
    L = m.nx
    rho = as2d(a["rho"], L, L)
    rho[:] = 0.5
    U = as2d(a["u"], L, L + 1)
    # cell j = 3 (row 3) sees faces 2, 3, 4, 5
    U[3, 2:6] = [0.0, 0.0, -1.0, -3.0]
    # row 4 expands: faces increasing
    U[4, :] = np.arange(L + 1, dtype=np.float64)
    po.lib().orc_artificial_viscosity(L, L, m.ref, 0.0, po.ptr(a["Qxx"]),
                                      po.ptr(a["Qyy"]), po.ptr(a["u"]),
                                      po.ptr(a["v"]), po.ptr(a["mom_x"]),
                                      po.ptr(a["mom_y"]), po.ptr(a["rho"]))
    Q = as2d(a["Qxx"], L, L)
    assert Q[3, 3] == 3.0 * 0.5
    assert np.all(Q[4, PAD:L - PAD] == 0.0)


def test_shock_heating_is_pdv_work(oracle):
    """Zero velocity leaves e unchanged bitwise; a uniform divergence D with
    Q = 0 changes e by -(p/rho) D dt_h to first order (de = -p dV / m)."""
    po = oracle
    n = 8
    m, a = grid_arrays(po, n)
    L = m.nx
    rho = as2d(a["rho"], L, L)
    e = as2d(a["e"], L, L)
    P = as2d(a["P"], L, L)
    rho[:] = 1.3
    e[:] = 2.0
    P[:] = 0.4 * 1.3 * 2.0
    e0 = e.copy()
    args = lambda dt_h: (L, L, m.ref, dt_h, po.ptr(a["e"]), po.ptr(a["P"]),
                         po.ptr(a["u"]), po.ptr(a["v"]), po.ptr(a["rho"]),
                         po.ptr(a["Qxx"]), po.ptr(a["Qyy"]))
    po.lib().orc_shock_heating_and_work(*args(1e-3))
    np.testing.assert_array_equal(e, e0)
    D = 0.7
    dx = 1.0 / n
    xf = (np.arange(L + 1) - PAD) * dx
    as2d(a["u"], L, L + 1)[:] = D * xf[None, :]
    for dt_h in (1e-4, 1e-5):
        e[:] = 2.0
        po.lib().orc_shock_heating_and_work(*args(dt_h))
        de = e[PAD:L - PAD, PAD:L - PAD] - 2.0
        first = -(0.4 * 1.3 * 2.0 / 1.3) * D * dt_h
        assert np.max(np.abs(de - first)) <= 5.0 * (D * dt_h) ** 2 * 2.0


def test_timestep_sound_speed_and_recurrence(oracle):
    """Uniform state, Q = 0: the CFL minimum is dx / c with the ideal-gas
    sound speed c^2 = gamma p / rho. The (dt, dt_h) update, worked by hand:
    dt_h = 1e-4 and a minimum far above it clamps to C_M dt_h = 5e-4, so
    dt = 0.5 (0.3 * 5e-4 + 1e-4) = 1.25e-4 and dt_h = 0.3 * 5e-4 = 1.5e-4;
    with first_step, dt_h = dt."""
    po = oracle
    n = 8
    m, a = grid_arrays(po, n, dt=1e-4 / 0.3)
    L = m.nx
    as2d(a["rho"], L, L)[:] = 1.7
    as2d(a["e"], L, L)[:] = 3.0
    m.m.dt, m.m.dt_h = 9.0, 1e-4
    g = po.set_timestep(m, a, 0)
    p = 0.4 * 1.7 * 3.0
    c = math.sqrt(1.4 * p / 1.7)
    assert abs(g - (1.0 / n) / c) <= 1e-15 * g
    assert m.m.dt == pytest.approx(1.25e-4, rel=1e-15)
    assert m.m.dt_h == pytest.approx(1.5e-4, rel=1e-15)
    m.m.dt, m.m.dt_h = 9.0, 1e-4
    po.set_timestep(m, a, 1)
    assert m.m.dt_h == m.m.dt
    # a small minimum is not clamped: dt = 0.5 (0.3 g + dt_h)
    m.m.dt_h = 1.0
    g = po.set_timestep(m, a, 0)
    assert m.m.dt == pytest.approx(0.5 * (0.3 * g + 1.0), rel=1e-15)
    assert m.m.dt_h == pytest.approx(0.3 * g, rel=1e-15)


# --------------------------------------------------------- advection stages ---

def _sweep_setup(po, n, rho_fn, e_fn, vel):
    m, a = grid_arrays(po, n)
    L = m.nx
    x = (np.arange(L) - PAD + 0.5) / n
    R, E = as2d(a["rho"], L, L), as2d(a["e"], L, L)
    R[:] = rho_fn(x)[None, :]
    E[:] = e_fn(x)[None, :]
    a["rho_old"][:] = a["rho"]
    as2d(a["u"], L, L + 1)[:] = vel
    return m, a, L


def _x_flux(po, m, a, L, dt, dt_h):
    po.lib().orc_x_mass_and_energy_flux(
        L, L, 1, m.ref, dt, dt_h, po.ptr(a["rho"]), po.ptr(a["rho_old"]),
        po.ptr(a["e"]), po.ptr(a["u"]), po.ptr(a["F_x"]), po.ptr(a["uF_x"]))
    return as2d(a["F_x"], L, L + 1), as2d(a["uF_x"], L, L + 1)


def test_x_flux_uniform_state(oracle):
    """Uniform rho, e and u: F = u rho and eF = dy e F exactly."""
    po = oracle
    n = 16
    m, a, L = _sweep_setup(po, n, lambda x: 0 * x + 1.25, lambda x: 0 * x + 2.0, 0.3)
    F, eF = _x_flux(po, m, a, L, 1e-3, 1e-3)
    inner = np.s_[PAD:L - PAD, PAD + 2:L - PAD - 2]
    assert np.all(F[inner] == 0.3 * 1.25)
    assert np.all(eF[inner] == (1.0 / n) * 2.0 * (0.3 * 1.25))


def test_x_flux_linear_density_is_lax_wendroff(oracle):
    """Linear density, uniform u > 0: the Van Leer ratio is 1, the limiter 1,
    and the flux is the Lax-Wendroff flux of linear advection,
    u rho_{j-1} + u/2 (1 - u dt/dx)(rho_j - rho_{j-1}). For u < 0 the upwind
    cell is j and the sign of the correction flips with |u|."""
    po = oracle
    n = 32
    dx = 1.0 / n
    for u in (0.4, -0.4):
        m, a, L = _sweep_setup(po, n, lambda x: 1.0 + 0.5 * x, lambda x: 0 * x + 1.0, u)
        dt_h = 1e-2
        R = as2d(a["rho"], L, L).copy()
        F, _ = _x_flux(po, m, a, L, 0.0, dt_h)
        i, j = 5, 10
        rm, rp = R[i, j - 1], R[i, j]
        up = rm if u >= 0 else rp
        lw = u * up + 0.5 * abs(u) * (1.0 - abs(u * dt_h / dx)) * (rp - rm)
        assert F[i, j] == pytest.approx(lw, rel=1e-14)


def test_x_flux_step_is_first_order_upwind(oracle):
    """At a density step the smoothness ratio is 0, the limiter vanishes and
    the flux is first-order upwind, u rho_upwind."""
    po = oracle
    n = 32
    m, a, L = _sweep_setup(po, n, lambda x: np.where(x < 0.5, 2.0, 1.0),
                           lambda x: 0 * x + 1.0, 0.25)
    F, _ = _x_flux(po, m, a, L, 0.0, 1e-2)
    j = PAD + n // 2  # face between the two states
    assert F[6, j] == 0.25 * 2.0


def test_x_flux_energy_face_value_linear(oracle):
    """Linear e with u > 0: the MC-limited slope is the exact slope s, and
    the face value is the Fromm/MUSCL reconstruction
    e_{j-1} + s/2 (dx - u dt_h); eF = dy e_face F."""
    po = oracle
    n = 32
    dx = 1.0 / n
    s = 3.0
    u = 0.2
    m, a, L = _sweep_setup(po, n, lambda x: 0 * x + 1.0, lambda x: 1.0 + s * x, u)
    dt_h = 1e-2
    E = as2d(a["e"], L, L).copy()
    F, eF = _x_flux(po, m, a, L, 0.0, dt_h)
    i, j = 7, 12
    e_face = E[i, j - 1] + 0.5 * s * (dx - u * dt_h)
    assert eF[i, j] == pytest.approx(dx * e_face * F[i, j], rel=1e-13)


def test_y_flux_mirrors_x_flux(oracle):
    """The y sweep on the transposed linear-density problem gives the same
    Lax-Wendroff flux (the Van Leer branch on v instead of v_tc only differs
    when the two have opposite signs)."""
    po = oracle
    n = 32
    dx = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    y = (np.arange(L) - PAD + 0.5) / n
    as2d(a["rho"], L, L)[:] = (1.0 + 0.5 * y)[:, None]
    as2d(a["e"], L, L)[:] = 1.0
    a["rho_old"][:] = a["rho"]
    as2d(a["v"], L + 1, L)[:] = 0.4
    dt_h = 1e-2
    R = as2d(a["rho"], L, L).copy()
    po.lib().orc_y_mass_and_energy_flux(
        L, L, 1, m.ref, 0.0, dt_h, po.ptr(a["rho"]), po.ptr(a["rho_old"]),
        po.ptr(a["e"]), po.ptr(a["v"]), po.ptr(a["F_y"]), po.ptr(a["uF_y"]))
    F = as2d(a["F_y"], L + 1, L)
    i, j = 10, 5
    rm, rp = R[i - 1, j], R[i, j]
    lw = 0.4 * rm + 0.2 * (1.0 - 0.4 * dt_h / dx) * (rp - rm)
    assert F[i, j] == pytest.approx(lw, rel=1e-14)


def test_momentum_flux_uniform_and_linear(oracle):
    """uF_x with uniform mass flux F and uniform u: F dy u exactly. With
    linear u (u > 0) the MC reconstruction is exact:
    u_j + s/2 (dx - u_cell dt_h)."""
    po = oracle
    n = 16
    dx = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    as2d(a["F_x"], L, L + 1)[:] = 0.8
    as2d(a["u"], L, L + 1)[:] = 0.5
    po.lib().orc_momentum_x_flux_in_x(L, L, m.ref, 1e-3, po.ptr(a["u"]),
                                      po.ptr(a["uF_x"]), po.ptr(a["F_x"]))
    uF = as2d(a["uF_x"], L, L)[PAD:L - PAD, PAD:L - PAD]
    assert np.all(uF == dx * 0.5 * (0.8 + 0.8) * 0.5)
    s, dt_h = 2.0, 1e-2
    xf = (np.arange(L + 1) - PAD) * dx
    U = as2d(a["u"], L, L + 1)
    U[:] = 0.3 + s * xf[None, :]
    po.lib().orc_momentum_x_flux_in_x(L, L, m.ref, dt_h, po.ptr(a["u"]),
                                      po.ptr(a["uF_x"]), po.ptr(a["F_x"]))
    i, j = 6, 7
    u_cell = 0.5 * (U[i, j] + U[i, j + 1])
    want = (dx * 0.8) * (U[i, j] + 0.5 * s * (dx - u_cell * dt_h))
    assert as2d(a["uF_x"], L, L)[i, j] == pytest.approx(want, rel=1e-13)


def _mom_flux(po, fn, m, dt_h, *arrs):
    getattr(po.lib(), fn)(m.nx, m.ny, m.ref, dt_h, *[po.ptr(x) for x in arrs])


def test_momentum_x_flux_in_y_closed_form(oracle):
    """uF_y at corner (i, j), omp3/flow.c:826-874. Uniform u: F u exactly,
    with F = dx (F_y[i, j] + F_y[i, j-1]) / 2 (F_y varies along x, so a
    shifted average fails). u linear in y with slope s: the MC slope is s,
    and the corner value is the upwind face value moved by half a cell
    against the corner velocity v_cell = (v[i, j-1] + v[i, j]) / 2: for
    v_cell > 0, u[i-1] + s/2 (dy - v_cell dt_h); otherwise u[i] - s/2 (dy +
    v_cell dt_h). v varies along x and changes sign, so both branches and
    the v_cell columns are tested."""
    po = oracle
    n = 16
    h = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    x = (np.arange(L) - PAD + 0.5) * h
    Fy = as2d(a["F_y"], L + 1, L)
    Fy[:] = (0.6 + 0.9 * x)[None, :]
    V = as2d(a["v"], L + 1, L)
    V[:] = (-0.45 + 1.1 * x)[None, :]
    U = as2d(a["u"], L, L + 1)
    U[:] = 0.35
    dt_h = 2e-2
    _mom_flux(po, "orc_momentum_x_flux_in_y", m, dt_h, a["u"], a["v"], a["uF_y"], a["F_y"])
    uF = as2d(a["uF_y"], L + 1, L + 1)
    for i, j in ((5, 3), (9, 14)):
        F = h * 0.5 * (Fy[i, j] + Fy[i, j - 1])
        assert uF[i, j] == pytest.approx(F * 0.35, rel=1e-14)
    s = 1.7
    yf = (np.arange(L) - PAD + 0.5) * h
    U[:] = (0.2 + s * yf)[:, None]
    _mom_flux(po, "orc_momentum_x_flux_in_y", m, dt_h, a["u"], a["v"], a["uF_y"], a["F_y"])
    for i, j in ((6, 3), (7, 4), (8, 13), (11, 15)):
        vc = 0.5 * (V[i, j - 1] + V[i, j])
        F = h * 0.5 * (Fy[i, j] + Fy[i, j - 1])
        if vc > 0.0:
            corner = U[i - 1, j] + 0.5 * s * (h - vc * dt_h)
        else:
            corner = U[i, j] - 0.5 * s * (h + vc * dt_h)
        assert uF[i, j] == pytest.approx(F * corner, rel=1e-13), (i, j, vc)
    assert (V[6, 2] + V[6, 3]) < 0 < (V[8, 12] + V[8, 13])


def test_momentum_y_flux_in_x_closed_form(oracle):
    """vF_x at corner (i, j), omp3/flow.c:877-925. Uniform v: F v exactly,
    F = dy (F_x[i, j] + F_x[i-1, j]) / 2 (F_x varies along y). v linear in x
    with slope s: for u_cell = (u[i, j] + u[i-1, j]) / 2 > 0 the value is
    v[j-1] + s/2 (dx - u_cell dt_h), otherwise v[j] - s/2 (dx + u_cell dt_h);
    u varies along y and changes sign."""
    po = oracle
    n = 16
    h = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    y = (np.arange(L) - PAD + 0.5) * h
    Fx = as2d(a["F_x"], L, L + 1)
    Fx[:] = (0.5 + 1.3 * y)[:, None]
    U = as2d(a["u"], L, L + 1)
    U[:] = (-0.5 + 1.2 * y)[:, None]
    V = as2d(a["v"], L + 1, L)
    V[:] = -0.25
    dt_h = 2e-2
    _mom_flux(po, "orc_momentum_y_flux_in_x", m, dt_h, a["u"], a["v"], a["vF_x"], a["F_x"])
    vF = as2d(a["vF_x"], L + 1, L + 1)
    for i, j in ((4, 6), (13, 9)):
        F = h * 0.5 * (Fx[i, j] + Fx[i - 1, j])
        assert vF[i, j] == pytest.approx(F * -0.25, rel=1e-14)
    s = -2.1
    xc = (np.arange(L) - PAD + 0.5) * h
    V[:] = (0.3 + s * xc)[None, :]
    _mom_flux(po, "orc_momentum_y_flux_in_x", m, dt_h, a["u"], a["v"], a["vF_x"], a["F_x"])
    for i, j in ((3, 5), (6, 8), (12, 7), (15, 10)):
        uc = 0.5 * (U[i, j] + U[i - 1, j])
        F = h * 0.5 * (Fx[i, j] + Fx[i - 1, j])
        if uc > 0.0:
            val = V[i, j - 1] + 0.5 * s * (h - uc * dt_h)
        else:
            val = V[i, j] - 0.5 * s * (h + uc * dt_h)
        assert vF[i, j] == pytest.approx(F * val, rel=1e-13), (i, j, uc)
    assert (U[3, 5] + U[2, 5]) < 0 < (U[12, 7] + U[11, 7])


def test_momentum_y_flux_in_y_closed_form(oracle):
    """vF_y at cell (i, j), omp3/flow.c:928-970. Uniform v: F v exactly,
    F = dx (F_y[i, j] + F_y[i+1, j]) / 2 (F_y varies along y). v linear in y
    with slope s: for v_cell = (v[i] + v[i+1]) / 2 > 0 the value is
    v[i] + s/2 (dy - v_cell dt_h), otherwise v[i+1] - s/2 (dy + v_cell dt_h)."""
    po = oracle
    n = 16
    h = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    yf = (np.arange(L + 1) - PAD) * h
    Fy = as2d(a["F_y"], L + 1, L)
    Fy[:] = (0.7 + 0.8 * yf)[:, None]
    V = as2d(a["v"], L + 1, L)
    V[:] = 0.45
    dt_h = 2e-2
    _mom_flux(po, "orc_momentum_y_flux_in_y", m, dt_h, a["v"], a["vF_y"], a["F_y"])
    vF = as2d(a["vF_y"], L, L)
    for i, j in ((5, 5), (12, 3)):
        F = h * 0.5 * (Fy[i, j] + Fy[i + 1, j])
        assert vF[i, j] == pytest.approx(F * 0.45, rel=1e-14)
    s = 1.4
    V[:] = (-0.6 + s * yf)[:, None]
    _mom_flux(po, "orc_momentum_y_flux_in_y", m, dt_h, a["v"], a["vF_y"], a["F_y"])
    for i, j in ((3, 4), (6, 9), (12, 10), (16, 3)):
        vc = 0.5 * (V[i, j] + V[i + 1, j])
        F = h * 0.5 * (Fy[i, j] + Fy[i + 1, j])
        if vc > 0.0:
            val = V[i, j] + 0.5 * s * (h - vc * dt_h)
        else:
            val = V[i + 1, j] - 0.5 * s * (h + vc * dt_h)
        assert vF[i, j] == pytest.approx(F * val, rel=1e-13), (i, j, vc)
    assert V[3, 4] + V[4, 4] < 0 < V[12, 10] + V[13, 10]


def test_y_flux_limiter_follows_v_not_v_tc(oracle):
    """omp3/flow.c:477: the y sweep picks the Van Leer ratio by the sign of
    the face velocity v, where the x sweep uses u_tc (:361). With v linear in
    y (slope 2) and dt = 2, v_tc = v (1 - dt s / 2) = -v: at a face with v > 0
    the upwind density is the upper cell (v_tc < 0) while the ratio is the
    lower one's, r = (rho[i-1] - rho[i-2]) / (rho[i] - rho[i-1]). Densities
    1.0, 1.1, 1.3, 1.9 make the lower and upper ratios 0.5 and 3."""
    po = oracle
    n = 16
    h = 1.0 / n
    m, a = grid_arrays(po, n)
    L = m.nx
    i = 8
    R = as2d(a["rho"], L, L)
    R[:] = 1.0
    R[i - 2], R[i - 1], R[i], R[i + 1] = 1.0, 1.1, 1.3, 1.9
    R[i + 2:] = 1.9
    as2d(a["e"], L, L)[:] = 1.0
    a["rho_old"][:] = a["rho"]
    yf = (np.arange(L + 1) - PAD) * h
    V = as2d(a["v"], L + 1, L)
    V[:] = (0.05 - 2.0 * yf[i] + 2.0 * yf)[:, None]   # v = 0.05 at face i
    dt, dt_h = 2.0, 1e-3
    po.lib().orc_y_mass_and_energy_flux(
        L, L, 1, m.ref, dt, dt_h, po.ptr(a["rho"]), po.ptr(a["rho_old"]),
        po.ptr(a["e"]), po.ptr(a["v"]), po.ptr(a["F_y"]), po.ptr(a["uF_y"]))
    F = as2d(a["F_y"], L + 1, L)
    v = V[i, 5]
    v_tc = v - 0.5 * v * dt * 2.0
    assert v > 0 > v_tc
    drho = 1.3 - 1.1
    r = (1.1 - 1.0) / drho
    lim = (r + abs(r)) / (1.0 + abs(r))
    want = v_tc * 1.3 + 0.5 * abs(v_tc) * (1.0 - abs(v_tc * dt_h / h)) * lim * drho
    assert F[i, 5] == pytest.approx(want, rel=1e-12)
    r_other = (1.9 - 1.3) / drho
    lim_other = (r_other + abs(r_other)) / (1.0 + abs(r_other))
    other = v_tc * 1.3 + 0.5 * abs(v_tc) * (1.0 - abs(v_tc * dt_h / h)) * lim_other * drho
    assert abs(other - want) > 1e-3 * abs(want)


# ------------------------------------------------------ whole-step invariants ---

def _run(po, n, steps, init):
    m = po.Mesh(n, n, dt=1e-4)
    a = po.new_arrays(m)
    a["rho"][:], a["e"][:] = init
    po.reflect_initial(m, a)
    out = po.run(m, a, steps)
    return m, a, out


def test_uniform_state_is_a_fixed_point(oracle):
    rho0, e0 = problems.uniform(16)
    m, a, (steps, dts, mass, energy, _) = _run(oracle, 16, 6, (rho0.copy(), e0.copy()))
    L = m.nx
    R = as2d(a["rho"], L, L)[PAD:L - PAD, PAD:L - PAD]
    E = as2d(a["e"], L, L)[PAD:L - PAD, PAD:L - PAD]
    assert np.all(R == 1.0) and np.all(E == 2.5)
    assert np.all(a["u"] == 0.0) and np.all(a["v"] == 0.0)


def test_sod_x_keeps_rows_identical_and_v_zero(oracle):
    m, a, _ = _run(oracle, 32, 8, problems.sod_x(32))
    L = m.nx
    R = as2d(a["rho"], L, L)[PAD:L - PAD, PAD:L - PAD]
    assert np.all(R == R[0][None, :])
    assert np.all(a["v"] == 0.0)
    assert np.any(as2d(a["u"], L, L + 1) != 0.0)


def test_mass_is_conserved(oracle):
    rho0, e0 = problems.multi_blast(48, seed=7)
    L = 48 + 2 * PAD
    m0 = math.fsum(as2d(rho0, L, L)[PAD:L - PAD, PAD:L - PAD].ravel())
    m, a, (steps, dts, mass, energy, _) = _run(oracle, 48, 12, (rho0, e0))
    assert np.max(np.abs(mass - m0)) <= 1e-13 * m0


def test_conservation_sums_brute_force(oracle):
    rho, e = problems.multi_blast(24, seed=3)
    L = 28
    ms, es = oracle.conservation(oracle.Mesh(24, 24), {"rho": rho, "e": e})
    R = as2d(rho, L, L)[PAD:L - PAD, PAD:L - PAD]
    E = as2d(e, L, L)[PAD:L - PAD, PAD:L - PAD]
    assert ms == pytest.approx(math.fsum(R.ravel()), rel=1e-14)
    assert es == pytest.approx(math.fsum((R * E).ravel()), rel=1e-14)


def test_corner_ghosts_do_not_reach_the_interior(oracle):
    """Corner ghost cells are never reflected; poisoning them with NaN must
    leave every interior value and the dt sequence unchanged."""
    init = problems.square_blast(24)
    m1, a1, o1 = _run(oracle, 24, 5, (init[0].copy(), init[1].copy()))
    rho, e = init[0].copy(), init[1].copy()
    L = 28
    for arr in (rho, e):
        A = as2d(arr, L, L)
        A[:PAD, :PAD] = A[:PAD, L - PAD:] = np.nan
        A[L - PAD:, :PAD] = A[L - PAD:, L - PAD:] = np.nan
    m2, a2, o2 = _run(oracle, 24, 5, (rho, e))
    for k in ("rho", "e"):
        A1 = as2d(a1[k], L, L)[PAD:L - PAD, PAD:L - PAD]
        A2 = as2d(a2[k], L, L)[PAD:L - PAD, PAD:L - PAD]
        np.testing.assert_array_equal(A1, A2)
    np.testing.assert_array_equal(o1[1], o2[1])


def test_rectangular_sod_y_keeps_columns_identical(oracle):
    """On an nx != ny mesh (12 x 40), a split along y keeps every column of
    rho identical and u zero, and conserves mass: a swapped stride or bound
    between x and y breaks all three."""
    from tests import harness
    nx, ny = 12, 40
    init = problems.boxes(nx, ny, [(0.125, 2.0, 0.0, 0.0, 1.0, 1.0),
                                   (1.0, 2.5, 0.0, 0.0, 1.0, 0.5)])
    Lx, Ly = nx + 2 * PAD, ny + 2 * PAD
    m0 = math.fsum(as2d(init[0], Ly, Lx)[PAD:Ly - PAD, PAD:Lx - PAD].ravel())
    r = harness.oracle_run(None, (nx, ny), 10, init=init)
    assert r["L"] == (Lx, Ly)
    R = as2d(r["arrays"]["rho"], Ly, Lx)[PAD:Ly - PAD, PAD:Lx - PAD]
    assert np.all(R == R[:, :1])
    assert np.any(R[:, 0] != R[0, 0])
    assert np.all(r["arrays"]["u"] == 0.0)
    assert np.any(as2d(r["arrays"]["v"], Ly + 1, Lx) != 0.0)
    assert np.max(np.abs(np.asarray(r["mass"]) - m0)) <= 1e-13 * m0
