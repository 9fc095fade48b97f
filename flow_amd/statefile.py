"""Bit-exact state files: the numpy side of flow_amd/csrc/hip/state.c.

A state file holds what one timestep carries to the next: rho and e (cell
layout, Ly x Lx), u and mom_x (x-face layout, Ly x (Lx+1)), v and mom_y
(y-face layout, (Ly+1) x Lx), with Lx = nx + 2*pad and Ly = ny + 2*pad
(rows x columns); the timestep pair (dt, dt_h) the next
step starts from; the steps completed and the simulated time. Step 0 is an
initial state, a later step a restart point. The driver (flow.hip --dump /
--load), the library (flowhip_state_save / _load) and this module read and
write the same bytes, so a workload generated here (python -m
flow_amd.statefile) starts the GPU run and the oracle from identical input.

Layout, little-endian: magic "FLOWSTAT", u32 version (1), u32 narrays (6),
i32 global_nx, global_ny, pad, step, f64 width, height, dt, dt_h, sim_time;
then per array (rho, e, u, v, mom_x, mom_y): name (16 bytes, NUL-padded),
layout ('C', 'X' or 'Y') and 7 reserved bytes, u64 count, count f64 values.
"""
import struct

import numpy as np

MAGIC = b"FLOWSTAT"
VERSION = 1
PAD = 2
ARRAYS = (("rho", "C"), ("e", "C"), ("u", "X"), ("v", "Y"), ("mom_x", "X"),
          ("mom_y", "Y"))
_HEADER = struct.Struct("<8sII4i5d")
_RECORD = struct.Struct("<16sc7sQ")


# Largest interior side the library indexes (FLOWHIP_MAX_CELLS_SIDE).
MAX_SIDE = 46335


def sides(n):
    """(nx, ny) interior cells of a mesh given as n (square) or (nx, ny)."""
    return (int(n), int(n)) if np.isscalar(n) else (int(n[0]), int(n[1]))


def layout_shape(layout, Lx, Ly=None):
    """Rows and columns of a layout for a mesh of Lx = nx + 2*pad columns of
    cells and Ly = ny + 2*pad rows (Ly defaults to Lx)."""
    Ly = Lx if Ly is None else Ly
    return {"C": (Ly, Lx), "X": (Ly, Lx + 1), "Y": (Ly + 1, Lx)}[layout]


def write(path, arrays, n, step=0, width=1.0, height=1.0, dt=0.0, dt_h=0.0,
          sim_time=0.0):
    """Write a state file for a mesh of n (square) or (nx, ny) interior
    cells. arrays maps each carried field name to either a 2-D array of its
    layout's shape or a flat buffer holding at least that many values
    (row-major, as the library's (Lx+1)(Ly+1) allocations do)."""
    nx, ny = sides(n)
    Lx, Ly = nx + 2 * PAD, ny + 2 * PAD
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, len(ARRAYS), nx, ny, PAD, int(step),
                             float(width), float(height), float(dt), float(dt_h),
                             float(sim_time)))
        for name, layout in ARRAYS:
            rows, cols = layout_shape(layout, Lx, Ly)
            a = np.ascontiguousarray(np.asarray(arrays[name], dtype="<f8").reshape(-1))
            if a.size < rows * cols:
                raise ValueError(f"{name}: {a.size} values, layout needs {rows * cols}")
            f.write(_RECORD.pack(name.encode(), layout.encode(), b"", rows * cols))
            f.write(a[: rows * cols].tobytes())


def read(path):
    """Read a state file: a dict with the header fields (n is nx) and
    'arrays', each a 2-D float64 array of its layout's shape."""
    with open(path, "rb") as f:
        head = f.read(_HEADER.size)
        if len(head) != _HEADER.size:
            raise ValueError(f"{path}: truncated header")
        (magic, version, narrays, nx, ny, pad, step, width, height, dt, dt_h,
         sim_time) = _HEADER.unpack(head)
        if magic != MAGIC:
            raise ValueError(f"{path}: not a state file")
        if (version != VERSION or narrays != len(ARRAYS) or pad != PAD
                or not (2 <= nx <= MAX_SIDE and 2 <= ny <= MAX_SIDE)):
            raise ValueError(f"{path}: unsupported state file")
        Lx, Ly = nx + 2 * pad, ny + 2 * pad
        arrays = {}
        for name, layout in ARRAYS:
            rec = f.read(_RECORD.size)
            if len(rec) != _RECORD.size:
                raise ValueError(f"{path}: truncated record")
            rname, rlayout, _, count = _RECORD.unpack(rec)
            rows, cols = layout_shape(layout, Lx, Ly)
            if (rname.rstrip(b"\0").decode() != name or rlayout.decode() != layout
                    or count != rows * cols):
                raise ValueError(f"{path}: unexpected record for {name}")
            data = np.frombuffer(f.read(8 * count), dtype="<f8")
            if data.size != count:
                raise ValueError(f"{path}: truncated data for {name}")
            arrays[name] = data.reshape(rows, cols).astype(np.float64)
    return dict(n=nx, nx=nx, ny=ny, pad=pad, step=step, width=width, height=height,
                dt=dt, dt_h=dt_h, sim_time=sim_time, arrays=arrays)


def initial_state(n, rho, e):
    """Carried fields of an initial state from flat (Lx+1)(Ly+1) rho and e
    buffers (the problems module's output) of a mesh of n or (nx, ny)
    interior cells: velocities and momenta zero."""
    nx, ny = sides(n)
    Lx, Ly = nx + 2 * PAD, ny + 2 * PAD
    out = {}
    for name, layout in ARRAYS:
        rows, cols = layout_shape(layout, Lx, Ly)
        src = {"rho": rho, "e": e}.get(name)
        out[name] = (np.asarray(src, dtype=np.float64)[: rows * cols].reshape(rows, cols)
                     if src is not None else np.zeros((rows, cols)))
    return out


def _main(argv=None):
    import argparse

    from flow_amd import problems

    p = argparse.ArgumentParser(
        description="Write a seeded workload's initial state file (step 0).")
    p.add_argument("problem", choices=("multi_blast", "blast_4096", "sod_x", "sod_y",
                                       "square_blast", "flow_params_problem"))
    p.add_argument("n", type=int)
    p.add_argument("out")
    p.add_argument("--seed", type=int, default=20250101,
                   help="multi_blast only (splitmix64 seed)")
    a = p.parse_args(argv)
    fn = getattr(problems, a.problem)
    rho, e = fn(a.n, seed=a.seed) if a.problem == "multi_blast" else fn(a.n)
    write(a.out, initial_state(a.n, rho, e), a.n)
    print(f"wrote {a.out}: {a.problem} n={a.n}")


if __name__ == "__main__":
    _main()
