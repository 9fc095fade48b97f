"""Initial states for the hydro timestep, as host arrays.

Each generator returns (rho, e) as flat float64 arrays of (L+1)^2 entries
(the allocation size of flow_data.c:8-15) whose first L*L entries hold the
cell layout i*L + j, L = n + 2*pad. Interior cells are set; ghost cells are
left zero for the initial reflection (main.c:56-59) to fill.

Box problems follow arch's rasterisation rule as main.c:52-54 uses it: a
cell takes a box's values when its left/bottom edge lies in
[pos, pos + extent), positions being fractions of the domain. Energies are
specific internal energies, e = p / ((gamma - 1) rho).
"""
import numpy as np

PAD = 2
GAMMA = 1.4


def local_size(n):
    return n + 2 * PAD


def edges(n, width=1.0):
    """edgex[k] for k in [0, L]: left edge of local cell k (uniform mesh)."""
    L = local_size(n)
    dx = width / n
    return dx * (np.arange(L + 1, dtype=np.float64) - PAD)


def _blank(nx, ny):
    L, M = local_size(nx), local_size(ny)
    n = (L + 1) * (M + 1)
    return np.zeros(n), np.zeros(n), L, M


def boxes(nx, ny, entries, width=1.0, height=1.0):
    """entries: [(rho, e, xpos, ypos, w, h), ...] applied in order, as the
    problem_N lines of a .params file (later entries overwrite)."""
    rho, e, L, M = _blank(nx, ny)
    ex, ey = edges(nx, width)[:L], edges(ny, height)[:M]
    R = rho[: L * M].reshape(M, L)
    E = e[: L * M].reshape(M, L)
    for (r, en, xp, yp, w, h) in entries:
        x0, y0, wx, hy = xp * width, yp * height, w * width, h * height
        cx = (ex >= x0) & (ex < x0 + wx)
        cy = (ey >= y0) & (ey < y0 + hy)
        mask = cy[:, None] & cx[None, :]
        R[mask] = r
        E[mask] = en
    return rho, e


def sod_x(n):
    """Sod-type split in x: x<0.5 rho=1 p=1 (e=2.5), else rho=0.125 p=0.1
    (e=2.0); the 512^2 CPU reference case uses n=512."""
    return boxes(n, n, [(0.125, 2.0, 0.0, 0.0, 1.0, 1.0),
                        (1.0, 2.5, 0.0, 0.0, 0.5, 1.0)])


def sod_y(n):
    """The Sod split along y."""
    return boxes(n, n, [(0.125, 2.0, 0.0, 0.0, 1.0, 1.0),
                        (1.0, 2.5, 0.0, 0.0, 1.0, 0.5)])


def square_blast(n):
    """Background rho=1 p=0.1; centre square of side 0.1 at p=100."""
    return boxes(n, n, [(1.0, 0.1 / (GAMMA - 1.0), 0.0, 0.0, 1.0, 1.0),
                        (1.0, 100.0 / (GAMMA - 1.0), 0.45, 0.45, 0.1, 0.1)])


def blast_4096(n):
    """The 4096^2 blast workload: rho=1, e=0.25 background; a centre box of
    side 0.1 at e=250 (409^2 cells at n=4096 by the left-edge rule)."""
    return boxes(n, n, [(1.0, 0.25, 0.0, 0.0, 1.0, 1.0),
                        (1.0, 250.0, 0.45, 0.45, 0.1, 0.1)])


def flow_params_problem(n):
    """The problem of the reference's flow.params (flow.params:1-2) on n x n
    interior cells, or nx x ny for n = (nx, ny)."""
    nx, ny = (n, n) if np.isscalar(n) else (int(n[0]), int(n[1]))
    return boxes(nx, ny, [(0.125, 2.0, 0.0, 0.0, 1.0, 1.0),
                          (1.0, 2.5, 0.4, 0.4, 0.2, 0.2)])


def uniform(n, rho0=1.0, e0=2.5):
    return boxes(n, n, [(rho0, e0, 0.0, 0.0, 1.0, 1.0)])


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): 64-bit state, golden-ratio
    increment, two xor-shift-multiply rounds. doubles() returns the top 53
    bits scaled to [0, 1)."""

    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * np.uint64(0x9E3779B97F4A7C15)
            self.state = self.state + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def doubles(self, n):
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def multi_blast(n, seed=20250101, side=0.05, grid=4, p_bg=0.1,
                p_lo=10.0, p_hi=1000.0, noise=0.05):
    """A grid x grid array of square blasts of the given side centred at
    ((k+1/2)/grid, (l+1/2)/grid), pressures log-uniform in [p_lo, p_hi], on a
    rho=1 p=p_bg background; every interior density is scaled by
    1 + U[-noise, noise]. Seeded splitmix64: first the grid*grid pressures
    (row-major, l then k), then one noise draw per interior cell (row-major).
    """
    rng = SplitMix64(seed)
    u = rng.doubles(grid * grid)
    pressures = np.exp(np.log(p_lo) + u * (np.log(p_hi) - np.log(p_lo)))
    L = local_size(n)
    rho = np.zeros((L + 1) * (L + 1))
    p = np.zeros_like(rho)
    ex = edges(n)[:L]
    R = rho[: L * L].reshape(L, L)
    Pm = p[: L * L].reshape(L, L)
    interior = np.s_[PAD:L - PAD, PAD:L - PAD]
    R[interior] = 1.0
    Pm[interior] = p_bg
    for l in range(grid):
        for k in range(grid):
            x0 = (k + 0.5) / grid - 0.5 * side
            y0 = (l + 0.5) / grid - 0.5 * side
            cx = (ex >= x0) & (ex < x0 + side)
            cy = (ex >= y0) & (ex < y0 + side)
            mask = cy[:, None] & cx[None, :]
            Pm[mask] = pressures[l * grid + k]
    w = rng.doubles(n * n).reshape(n, n)
    R[interior] = R[interior] * (1.0 + noise * (2.0 * w - 1.0))
    e = np.zeros_like(rho)
    E = e[: L * L].reshape(L, L)
    E[interior] = Pm[interior] / ((GAMMA - 1.0) * R[interior])
    return rho, e


def ensemble_specs(count=256, seed=20250102, n_lo=64, n_hi=1024,
                   split_lo=0.3, split_hi=0.7, ratio_lo=10.0, ratio_hi=1000.0):
    """Instances of the batching workload: per instance (seeded splitmix64,
    four draws each, in this order) a side n log-uniform in [n_lo, n_hi]
    (rounded down), a split position U[split_lo, split_hi], a pressure ratio
    log-uniform in [ratio_lo, ratio_hi] and a direction draw. The high side
    is rho=1, p=0.1*ratio, the low side rho=0.125, p=0.1 (a Sod-type shock
    tube); even instances split along x, odd ones along y unless the
    direction draw is below 0.5, which makes the high-pressure region a
    centred square blast of side 2*(split - 0.3) + 0.1 instead.
    Returns [(n, kind, split, ratio)]."""
    rng = SplitMix64(seed)
    u = rng.doubles(4 * count).reshape(count, 4)
    out = []
    for k in range(count):
        n = int(np.floor(np.exp(np.log(n_lo) + u[k, 0] * (np.log(n_hi) - np.log(n_lo)))))
        n = max(n_lo, min(n_hi, n))
        split = split_lo + u[k, 1] * (split_hi - split_lo)
        ratio = np.exp(np.log(ratio_lo) + u[k, 2] * (np.log(ratio_hi) - np.log(ratio_lo)))
        if u[k, 3] < 0.5:
            kind = "blast"
        else:
            kind = "sod_x" if k % 2 == 0 else "sod_y"
        out.append((n, kind, float(split), float(ratio)))
    return out


def ensemble_instance(n, kind, split, ratio):
    """(rho, e) of one ensemble instance (see ensemble_specs)."""
    e_hi = 0.1 * ratio / ((GAMMA - 1.0) * 1.0)
    e_lo = 0.1 / ((GAMMA - 1.0) * 0.125)
    if kind == "sod_x":
        return boxes(n, n, [(0.125, e_lo, 0.0, 0.0, 1.0, 1.0),
                            (1.0, e_hi, 0.0, 0.0, split, 1.0)])
    if kind == "sod_y":
        return boxes(n, n, [(0.125, e_lo, 0.0, 0.0, 1.0, 1.0),
                            (1.0, e_hi, 0.0, 0.0, 1.0, split)])
    side = 2.0 * (split - 0.3) + 0.1
    lo = 0.5 - 0.5 * side
    return boxes(n, n, [(1.0, 0.1 / (GAMMA - 1.0), 0.0, 0.0, 1.0, 1.0),
                        (1.0, e_hi, lo, lo, side, side)])


PROBLEMS = {
    "sod_x": sod_x,
    "sod_y": sod_y,
    "square_blast": square_blast,
    "flow_params": flow_params_problem,
    "uniform": uniform,
    "multi_blast": multi_blast,
}
