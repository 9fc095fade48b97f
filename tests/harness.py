"""Shared helpers for the parity tests: run one problem through the oracle
(CPU restatement of omp3) and through the hip kernel set on identical
inputs, and compare the results over the regions omp3 defines."""
import numpy as np

from flow_amd import problems

PAD = 2


def oracle_run(problem_fn, n, steps, width=1.0, max_dt=1.0e30, dt=1.0e-4,
               init=None):
    from oracle import pyoracle as po
    m = po.Mesh(n, n, width=width, height=width, max_dt=max_dt, dt=dt)
    a = po.new_arrays(m)
    rho, e = init if init is not None else problem_fn(n)
    a["rho"][:] = rho
    a["e"][:] = e
    po.reflect_initial(m, a)
    steps_done, dts, mass, energy, wall = po.run(m, a, steps)
    return dict(arrays=a, dt=dts, mass=mass, energy=energy, steps=steps_done,
                L=m.nx, wall=wall)


def gpu_run(problem_fn, n, steps, schedule, width=1.0, max_dt=1.0e30,
            dt=1.0e-4, graphs=False, init=None, march_rows=None):
    import flow_amd
    m = flow_amd.Mesh(n, n, width=width, height=width, max_dt=max_dt, dt=dt)
    m.set_schedule(schedule)
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
               steps=steps_done, L=m.nx, wall=wall)
    st.close()
    m.close()
    return out


def views(a, L):
    """2-D views of the flat arrays by layout."""
    return {
        "rho": a["rho"][: L * L].reshape(L, L),
        "e": a["e"][: L * L].reshape(L, L),
        "u": a["u"][: L * (L + 1)].reshape(L, L + 1),
        "v": a["v"][: (L + 1) * L].reshape(L + 1, L),
        "mom_x": a["mom_x"][: L * (L + 1)].reshape(L, L + 1),
        "mom_y": a["mom_y"][: (L + 1) * L].reshape(L + 1, L),
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
    (omp3's FX / FY, omp3/flow.c:578-580, 631-633)."""
    mx = np.zeros((L, L + 1), dtype=bool)
    mx[PAD:L - PAD, PAD:L + 1 - PAD] = True
    my = np.zeros((L + 1, L), dtype=bool)
    my[PAD:L + 1 - PAD, PAD:L - PAD] = True
    return {
        "rho": noncorner_mask(L, L), "e": noncorner_mask(L, L),
        "u": noncorner_mask(L, L + 1), "v": noncorner_mask(L + 1, L),
        "mom_x": mx, "mom_y": my,
    }


def rel_linf(a, b):
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0.0:
        return float(num)
    return float(num / den)


def compare_state(got, ref, L, fields=("rho", "e", "u", "v", "mom_x", "mom_y")):
    """Per field: (relative L-inf over the defined region, number of
    elements that differ bitwise)."""
    g, r = views(got, L), views(ref, L)
    reg = state_regions(L)
    out = {}
    for k in fields:
        a, b = g[k][reg[k]], r[k][reg[k]]
        out[k] = (rel_linf(a, b), int(np.count_nonzero(a != b)))
    return out
