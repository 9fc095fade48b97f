"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: a CPU restatement of UoB-HPC/flow's omp3 hydro
timestep. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module, and only as the checker or the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NO_INVERT, INVERT_X, INVERT_Y = 0, 1, 2

# Order of the 16 arrays solve_hydro_2d takes (omp3/flow.c:10-16).
ARRAY_NAMES = ("P", "rho", "rho_old", "e", "u", "v", "mom_x", "mom_y", "Qxx",
               "Qyy", "F_x", "F_y", "uF_x", "uF_y", "vF_x", "vF_y")

_dp = ctypes.POINTER(ctypes.c_double)


class OrcMesh(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("pad", ctypes.c_int),
                ("global_nx", ctypes.c_int), ("global_ny", ctypes.c_int),
                ("width", ctypes.c_double), ("height", ctypes.c_double),
                ("max_dt", ctypes.c_double), ("sim_end", ctypes.c_double),
                ("dt", ctypes.c_double), ("dt_h", ctypes.c_double),
                ("edgex", _dp), ("edgey", _dp), ("edgedx", _dp), ("edgedy", _dp),
                ("celldx", _dp), ("celldy", _dp), ("verbose", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        mp = ctypes.POINTER(OrcMesh)
        L.orc_minmod.restype = d
        L.orc_minmod.argtypes = [d, d]
        L.orc_min.restype = d
        L.orc_min.argtypes = [d, d]
        L.orc_mesh_init.argtypes = [mp, i, i, d, d, d, d, d]
        L.orc_mesh_free.argtypes = [mp]
        L.orc_handle_boundary_2d.argtypes = [i, i, i, vp, i]
        L.orc_rasterise_box.argtypes = [mp, vp, vp, d, d, d, d, d, d]
        L.orc_equation_of_state.argtypes = [i, i, vp, vp, vp]
        L.orc_pressure_acceleration.argtypes = [i, i, mp, d] + [vp] * 6
        L.orc_artificial_viscosity.argtypes = [i, i, mp, d] + [vp] * 7
        L.orc_shock_heating_and_work.argtypes = [i, i, mp, d] + [vp] * 7
        L.orc_set_timestep.restype = d
        L.orc_set_timestep.argtypes = [i, i, vp, vp, vp, vp, mp, i]
        L.orc_set_timestep_literal.restype = d
        L.orc_set_timestep_literal.argtypes = [i, i, vp, vp, vp, vp, mp, i, i]
        L.orc_set_cfl_rule.argtypes = [i]
        L.orc_get_cfl_rule.restype = i
        L.orc_last_undefined.restype = i
        L.orc_x_mass_and_energy_flux.argtypes = [i, i, i, mp, d, d] + [vp] * 6
        L.orc_y_mass_and_energy_flux.argtypes = [i, i, i, mp, d, d] + [vp] * 6
        L.orc_advect_mass_and_energy.argtypes = [i, i, mp, i, d, d] + [vp] * 9
        L.orc_momentum_x_flux_in_x.argtypes = [i, i, mp, d, vp, vp, vp]
        L.orc_momentum_x_flux_in_y.argtypes = [i, i, mp, d, vp, vp, vp, vp]
        L.orc_momentum_y_flux_in_x.argtypes = [i, i, mp, d, vp, vp, vp, vp]
        L.orc_momentum_y_flux_in_y.argtypes = [i, i, mp, d, vp, vp, vp]
        L.orc_advect_momentum.argtypes = [i, i, i, mp, d, d] + [vp] * 11
        L.orc_conservation.argtypes = [i, i, i, vp, vp, _dp, _dp]
        L.orc_solve_hydro_2d.argtypes = [mp, i] + [vp] * 16
        L.orc_run.restype = i
        L.orc_run.argtypes = [mp, ctypes.POINTER(vp), i, vp, vp, vp, _dp, vp]
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """Uniform mesh + timestep pair, the oracle's restatement of arch's Mesh."""

    def __init__(self, global_nx, global_ny, width=1.0, height=1.0,
                 max_dt=1.0e30, sim_end=1.0e30, dt=1.0e-4):
        self.m = OrcMesh()
        lib().orc_mesh_init(ctypes.byref(self.m), global_nx, global_ny, width,
                            height, max_dt, sim_end, dt)

    def __del__(self):
        try:
            lib().orc_mesh_free(ctypes.byref(self.m))
        except Exception:
            pass

    @property
    def ref(self):
        return ctypes.byref(self.m)

    @property
    def nx(self):
        return self.m.nx

    @property
    def ny(self):
        return self.m.ny

    @property
    def pad(self):
        return self.m.pad

    def geometry(self, name):
        n = (self.m.nx if name.endswith("x") else self.m.ny) + 1
        return np.ctypeslib.as_array(getattr(self.m, name), shape=(n,)).copy()


def new_arrays(mesh):
    """The 16 solve_hydro_2d arrays, each (L+1)^2 zeros (flow_data.c:8-15)."""
    n = (mesh.nx + 1) * (mesh.ny + 1)
    return {k: np.zeros(n, dtype=np.float64) for k in ARRAY_NAMES}


def boundary(nx, ny, pad, arr, invert):
    lib().orc_handle_boundary_2d(nx, ny, pad, ptr(arr), invert)


def rasterise_box(mesh, arrays, rho, e, xpos, ypos, w, h):
    lib().orc_rasterise_box(mesh.ref, ptr(arrays["rho"]), ptr(arrays["e"]),
                            rho, e, xpos, ypos, w, h)


def reflect_initial(mesh, arrays):
    """main.c:56-59: reflect density and energy before the first step."""
    boundary(mesh.nx, mesh.ny, mesh.pad, arrays["rho"], NO_INVERT)
    boundary(mesh.nx, mesh.ny, mesh.pad, arrays["e"], NO_INVERT)


def solve_step(mesh, arrays, tt):
    lib().orc_solve_hydro_2d(mesh.ref, tt, *[ptr(arrays[k]) for k in ARRAY_NAMES])


def set_timestep(mesh, arrays, first_step):
    return lib().orc_set_timestep(mesh.nx, mesh.ny, ptr(arrays["Qxx"]),
                                  ptr(arrays["Qyy"]), ptr(arrays["rho"]),
                                  ptr(arrays["e"]), mesh.ref, first_step)


# Rules for NaN CFL candidates (oracle.h): the minimum over the defined
# candidates, or omp3's fold as one OpenMP thread computes it.
CFL_DEFINED, CFL_OMP3_SERIAL = 0, 1


def set_cfl_rule(rule):
    lib().orc_set_cfl_rule(rule)


def get_cfl_rule():
    return lib().orc_get_cfl_rule()


class cfl_rule:
    """Context manager: run the oracle under one CFL rule, then restore."""

    def __init__(self, rule):
        self.rule = rule

    def __enter__(self):
        self.prev = get_cfl_rule()
        set_cfl_rule(self.rule)
        return self

    def __exit__(self, *exc):
        set_cfl_rule(self.prev)


def set_timestep_literal(mesh, arrays, first_step, nthreads):
    """omp3's loop form (a min reduction over nthreads threads); test only."""
    return lib().orc_set_timestep_literal(
        mesh.nx, mesh.ny, ptr(arrays["Qxx"]), ptr(arrays["Qyy"]),
        ptr(arrays["rho"]), ptr(arrays["e"]), mesh.ref, first_step, nthreads)


def conservation(mesh, arrays):
    m, e = ctypes.c_double(), ctypes.c_double()
    lib().orc_conservation(mesh.nx, mesh.ny, mesh.pad, ptr(arrays["rho"]),
                           ptr(arrays["e"]), ctypes.byref(m), ctypes.byref(e))
    return m.value, e.value


def run(mesh, arrays, niters, undefined=None):
    """main.c:64-115 without I/O. Returns (steps, dt_pairs, mass, energy,
    wall_s). undefined, if given, is an int32 array of niters + 1 that
    receives the NaN CFL candidates of the pre-loop call and of each step."""
    dt_seq = np.zeros(2 * (niters + 1))
    mass = np.zeros(niters)
    energy = np.zeros(niters)
    wall = ctypes.c_double()
    ptrs = (ctypes.c_void_p * 16)(*[arrays[k].ctypes.data for k in ARRAY_NAMES])
    if undefined is not None:
        assert undefined.dtype == np.int32 and undefined.size >= niters + 1
        und = undefined.ctypes.data_as(ctypes.c_void_p)
    else:
        und = None
    steps = lib().orc_run(mesh.ref, ptrs, niters, ptr(dt_seq), ptr(mass),
                          ptr(energy), ctypes.byref(wall), und)
    return (steps, dt_seq[: 2 * (steps + 1)].reshape(-1, 2), mass[:steps],
            energy[:steps], wall.value)
