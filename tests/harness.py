"""Shared helpers for the parity tests: run one problem through the oracle
(CPU restatement of omp3) and through the hip kernel set on identical
inputs, and compare the results over the regions omp3 defines."""
import numpy as np

from flow_amd import problems

PAD = 2


def sides(n):
    """(nx, ny) of a mesh given as n (square) or (nx, ny)."""
    return (n, n) if np.isscalar(n) else (int(n[0]), int(n[1]))


def oracle_run(problem_fn, n, steps, width=1.0, max_dt=1.0e30, dt=1.0e-4,
               init=None, cfl_rule=0):
    from oracle import pyoracle as po
    nx, ny = sides(n)
    m = po.Mesh(nx, ny, width=width, height=width, max_dt=max_dt, dt=dt)
    a = po.new_arrays(m)
    rho, e = init if init is not None else problem_fn(n)
    a["rho"][:] = rho
    a["e"][:] = e
    po.reflect_initial(m, a)
    und = np.zeros(steps + 1, dtype=np.int32)
    with po.cfl_rule(cfl_rule):
        steps_done, dts, mass, energy, wall = po.run(m, a, steps, undefined=und)
    return dict(arrays=a, dt=dts, mass=mass, energy=energy, steps=steps_done,
                L=m.nx if nx == ny else (m.nx, m.ny), wall=wall,
                undefined=und[: steps_done + 1])


def gpu_run(problem_fn, n, steps, schedule, width=1.0, max_dt=1.0e30,
            dt=1.0e-4, graphs=False, init=None, march_rows=None, cfl_rule=0):
    import flow_amd
    nx, ny = sides(n)
    m = flow_amd.Mesh(nx, ny, width=width, height=width, max_dt=max_dt, dt=dt)
    m.set_schedule(schedule)
    m.set_cfl_rule(cfl_rule)
    if march_rows:
        m.set_march_rows(march_rows)
    m.set_graphs(graphs)
    st = flow_amd.DeviceState(m)
    rho, e = init if init is not None else problem_fn(n)
    st.upload("rho", rho)
    st.upload("e", e)
    st.reflect_initial()
    steps_done, dts, mass, energy, wall = st.run(steps)
    out = dict(arrays=st.download_all(), dt=dts, mass=mass, energy=energy,
               steps=steps_done, L=m.nx if nx == ny else (m.nx, m.ny), wall=wall,
               undefined=st.undefined)
    st.close()
    m.close()
    return out


def views(a, L):
    """2-D views of the flat arrays by layout; L is the padded side, or
    (Lx, Ly). Cell arrays are Ly rows of Lx, x-face arrays Ly rows of Lx+1,
    y-face arrays Ly+1 rows of Lx (strides local_nx and local_nx + 1)."""
    Lx, Ly = sides(L)
    return {
        "rho": a["rho"][: Lx * Ly].reshape(Ly, Lx),
        "e": a["e"][: Lx * Ly].reshape(Ly, Lx),
        "u": a["u"][: Ly * (Lx + 1)].reshape(Ly, Lx + 1),
        "v": a["v"][: (Ly + 1) * Lx].reshape(Ly + 1, Lx),
        "mom_x": a["mom_x"][: Ly * (Lx + 1)].reshape(Ly, Lx + 1),
        "mom_y": a["mom_y"][: (Ly + 1) * Lx].reshape(Ly + 1, Lx),
    }


def noncorner_mask(rows, cols, pad=PAD):
    """True except in the four pad x pad corner blocks (which the reflective
    boundary never writes)."""
    m = np.ones((rows, cols), dtype=bool)
    m[:pad, :pad] = m[:pad, cols - pad:] = False
    m[rows - pad:, :pad] = m[rows - pad:, cols - pad:] = False
    return m


def state_regions(L):
    """Where each carried field is defined after a step: cell fields and
    velocities everywhere but the corners; momenta on their face domains
    (omp3's FX / FY, omp3/flow.c:578-580, 631-633). L as for views."""
    Lx, Ly = sides(L)
    mx = np.zeros((Ly, Lx + 1), dtype=bool)
    mx[PAD:Ly - PAD, PAD:Lx + 1 - PAD] = True
    my = np.zeros((Ly + 1, Lx), dtype=bool)
    my[PAD:Ly + 1 - PAD, PAD:Lx - PAD] = True
    return {
        "rho": noncorner_mask(Ly, Lx), "e": noncorner_mask(Ly, Lx),
        "u": noncorner_mask(Ly, Lx + 1), "v": noncorner_mask(Ly + 1, Lx),
        "mom_x": mx, "mom_y": my,
    }


def rel_linf(a, b):
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0.0:
        return float(num)
    return float(num / den)


def compare_state(got, ref, L, fields=("rho", "e", "u", "v", "mom_x", "mom_y"),
                  nan_equal=False):
    """Per field: (relative L-inf over the defined region, number of
    elements that differ bitwise). With nan_equal, a NaN on both sides
    counts as equal (NaN payloads are not compared) and the L-inf is taken
    over the elements finite on both sides."""
    g, r = views(got, L), views(ref, L)
    reg = state_regions(L)
    out = {}
    for k in fields:
        a, b = g[k][reg[k]], r[k][reg[k]]
        if nan_equal:
            both = np.isnan(a) & np.isnan(b)
            ndiff = int(np.count_nonzero((a != b) & ~both))
            fin = np.isfinite(a) & np.isfinite(b)
            rel = rel_linf(a[fin], b[fin]) if fin.any() else 0.0
            out[k] = (rel, ndiff)
        else:
            out[k] = (rel_linf(a, b), int(np.count_nonzero(a != b)))
    return out
