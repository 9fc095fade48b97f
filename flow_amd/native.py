"""ctypes binding of libflowhip.so (include/flow_interface.h, flow_arch.h,
flow_hip.h). The library is the only compute path: if it is missing this
module raises instead of falling back to anything else."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.environ.get("FLOWHIP_LIBRARY") or os.path.join(_HERE, "lib", "libflowhip.so")

STAGEWISE, FUSED, MARCH = 0, 1, 2
# Rules for undefined CFL candidates (flow_hip.h FLOWHIP_CFL_*).
CFL_DEFINED, CFL_OMP3_SERIAL = 0, 1
NO_INVERT, INVERT_X, INVERT_Y = 0, 1, 2
RECV, SEND = 0, 1

# solve_hydro_2d's array arguments in order (flow_interface.h:16-20).
ARRAY_NAMES = ("P", "rho", "rho_old", "e", "u", "v", "mom_x", "mom_y", "Qxx",
               "Qyy", "F_x", "F_y", "uF_x", "uF_y", "vF_x", "vF_y")

_dp = ctypes.POINTER(ctypes.c_double)


class MeshStruct(ctypes.Structure):
    """Mirror of `Mesh` in include/flow_arch.h."""
    _fields_ = [("local_nx", ctypes.c_int), ("local_ny", ctypes.c_int),
                ("global_nx", ctypes.c_int), ("global_ny", ctypes.c_int),
                ("niters", ctypes.c_int), ("pad", ctypes.c_int),
                ("rank", ctypes.c_int), ("nranks", ctypes.c_int),
                ("x_off", ctypes.c_int), ("y_off", ctypes.c_int),
                ("neighbours", ctypes.c_int * 4),
                ("width", ctypes.c_double), ("height", ctypes.c_double),
                ("max_dt", ctypes.c_double), ("sim_end", ctypes.c_double),
                ("dt", ctypes.c_double), ("dt_h", ctypes.c_double),
                ("edgex", _dp), ("edgey", _dp), ("edgedx", _dp),
                ("edgedy", _dp), ("celldx", _dp), ("celldy", _dp),
                ("verbose", ctypes.c_int), ("hip", ctypes.c_void_p)]


_lib = None


def library_path():
    return _LIB


def lib():
    """Load libflowhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        raise RuntimeError(
            f"hip kernel set not built: {_LIB} is missing "
            "(run __graft_entry__.build() or make -C flow_amd/csrc)")
    L = ctypes.CDLL(_LIB)
    i, d, vp, sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
    mp = ctypes.POINTER(MeshStruct)
    sig = {
        "flowhip_mesh_new": (mp, [i, i, d, d, d, d, d]),
        "flowhip_mesh_delete": (None, [mp]),
        "flowhip_set_schedule": (i, [mp, i]),
        "flowhip_get_schedule": (i, [mp]),
        "flowhip_set_graphs": (i, [mp, i]),
        "flowhip_set_march_rows": (i, [mp, i]),
        "flowhip_set_timing": (i, [mp, i]),
        "flowhip_timing_count": (i, [mp]),
        "flowhip_timing_entry": (i, [mp, i, ctypes.c_char_p, i, _dp,
                                     ctypes.POINTER(ctypes.c_long)]),
        "flowhip_timing_reset": (i, [mp]),
        "flowhip_run": (i, [mp, ctypes.POINTER(vp), i, i, i, vp, vp, vp, _dp, vp]),
        "flowhip_solve_hydro_2d_static": (i, [mp, i] + [vp] * 16 + [_dp, _dp, vp]),
        "flowhip_set_cfl_rule": (i, [mp, i]),
        "flowhip_get_cfl_rule": (i, [mp]),
        "flowhip_undefined_candidates": (i, [mp]),
        "flowhip_device_sync": (i, []),
        "flowhip_memset_zero": (i, [vp, sz]),
        "flowhip_copy": (i, [vp, vp, sz, i]),
        "flowhip_device_name": (i, [ctypes.c_char_p, i]),
        "flowhip_f32_alloc": (i, [ctypes.POINTER(vp), sz]),
        "flowhip_f32_free": (None, [vp]),
        "flowhip_f32_zero": (i, [vp, sz]),
        "flowhip_f32_copy": (i, [vp, vp, sz, i]),
        "flowhip_f32_narrow": (i, [vp, vp, sz]),
        "flowhip_f32_widen": (i, [vp, vp, sz]),
        "flowhip_f32_boundary": (i, [mp, vp, i, i, i]),
        "flowhip_run_f32": (i, [mp, ctypes.POINTER(vp), i, i, i, vp, vp, vp, _dp, vp]),
        "flowhip_ensemble_new": (vp, [i, ctypes.POINTER(i), ctypes.POINTER(i), _dp, _dp, _dp, _dp, i]),
        "flowhip_ensemble_delete": (None, [vp]),
        "flowhip_ensemble_size": (i, [vp]),
        "flowhip_ensemble_local_size": (i, [vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "flowhip_ensemble_array": (vp, [vp, i, i]),
        "flowhip_ensemble_set_state": (i, [vp, i, vp, vp]),
        "flowhip_ensemble_get_array": (i, [vp, i, i, vp]),
        "flowhip_ensemble_set_dt": (i, [vp, i, d, d]),
        "flowhip_ensemble_set_graphs": (i, [vp, i]),
        "flowhip_ensemble_set_schedule": (i, [vp, i, i, i]),
        "flowhip_ensemble_get_schedule": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "flowhip_ensemble_autotune": (i, [vp, i, i, _dp]),
        "flowhip_ensemble_run": (i, [vp, i, i, i, vp, vp, vp, _dp, vp]),
        "allocate_data": (sz, [ctypes.POINTER(vp), sz]),
        "deallocate_data": (None, [vp]),
        "handle_boundary_2d": (None, [i, i, mp, vp, i, i]),
        "solve_hydro_2d": (None, [mp, i] + [vp] * 17),
        "set_timestep": (None, [i, i, vp, vp, vp, vp, mp, vp, i, vp, vp]),
        "print_conservation": (None, [i, i, vp, vp, vp, mp]),
        "get_int_parameter": (i, [ctypes.c_char_p, ctypes.c_char_p]),
        "get_double_parameter": (d, [ctypes.c_char_p, ctypes.c_char_p]),
        "get_key_value_parameter": (i, [ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.c_char_p, _dp,
                                        ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """A Mesh from flowhip_mesh_new (main.c:20-40 with arch.params values)."""

    def __init__(self, global_nx, global_ny, width=1.0, height=1.0,
                 max_dt=1.0e30, sim_end=1.0e30, dt=1.0e-4, verbose=False):
        self.ptr = lib().flowhip_mesh_new(global_nx, global_ny, width, height,
                                          max_dt, sim_end, dt)
        if not self.ptr:
            raise RuntimeError("flowhip_mesh_new failed")
        self.ptr.contents.verbose = 1 if verbose else 0

    def close(self):
        if self.ptr:
            lib().flowhip_mesh_delete(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def s(self):
        return self.ptr.contents

    nx = property(lambda self: self.s.local_nx)
    ny = property(lambda self: self.s.local_ny)
    pad = property(lambda self: self.s.pad)
    dt = property(lambda self: self.s.dt)
    dt_h = property(lambda self: self.s.dt_h)

    def set_dt(self, dt, dt_h):
        self.s.dt = dt
        self.s.dt_h = dt_h

    def set_schedule(self, schedule):
        if lib().flowhip_set_schedule(self.ptr, schedule) != 0:
            raise ValueError(f"bad schedule {schedule}")

    def set_march_rows(self, rows):
        if lib().flowhip_set_march_rows(self.ptr, rows) != 0:
            raise ValueError(f"bad chunk height {rows}")

    def set_graphs(self, on):
        lib().flowhip_set_graphs(self.ptr, 1 if on else 0)

    def set_cfl_rule(self, rule):
        """CFL_DEFINED or CFL_OMP3_SERIAL (flow_hip.h)."""
        if lib().flowhip_set_cfl_rule(self.ptr, rule) != 0:
            raise ValueError(f"bad CFL rule {rule}")

    def cfl_rule(self):
        return lib().flowhip_get_cfl_rule(self.ptr)

    def undefined_candidates(self):
        """Undefined CFL candidates met by the last dt computation."""
        return lib().flowhip_undefined_candidates(self.ptr)

    def set_timing(self, on):
        lib().flowhip_set_timing(self.ptr, 1 if on else 0)

    def timing(self):
        out = {}
        name = ctypes.create_string_buffer(128)
        ms, n = ctypes.c_double(), ctypes.c_long()
        for k in range(lib().flowhip_timing_count(self.ptr)):
            lib().flowhip_timing_entry(self.ptr, k, name, 128, ctypes.byref(ms),
                                       ctypes.byref(n))
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def reset_timing(self):
        lib().flowhip_timing_reset(self.ptr)


class DeviceState:
    """The 16 device arrays solve_hydro_2d takes, each (L+1)^2 doubles
    allocated zeroed by allocate_data (flow_data.c:8-15)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.n = (mesh.nx + 1) * (mesh.ny + 1)
        self.ptrs = {}
        for k in ARRAY_NAMES:
            p = ctypes.c_void_p()
            lib().allocate_data(ctypes.byref(p), self.n)
            self.ptrs[k] = p.value

    def close(self):
        for k, p in list(self.ptrs.items()):
            lib().deallocate_data(ctypes.c_void_p(p))
        self.ptrs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, name, host):
        host = np.ascontiguousarray(host, dtype=np.float64).ravel()
        buf = np.zeros(self.n, dtype=np.float64)
        buf[: host.size] = host
        lib().flowhip_copy(ctypes.c_void_p(self.ptrs[name]), _ptr(buf), self.n, SEND)

    def download(self, name):
        buf = np.empty(self.n, dtype=np.float64)
        lib().flowhip_copy(ctypes.c_void_p(self.ptrs[name]), _ptr(buf), self.n, RECV)
        return buf

    def download_all(self):
        return {k: self.download(k) for k in ARRAY_NAMES}

    def zero(self, name):
        lib().flowhip_memset_zero(ctypes.c_void_p(self.ptrs[name]), self.n)

    def boundary(self, name, nxp, nyp, invert):
        lib().handle_boundary_2d(nxp, nyp, self.mesh.ptr,
                                 ctypes.c_void_p(self.ptrs[name]), invert, 1)

    def reflect_initial(self):
        """main.c:56-59: reflect density and energy before the first step."""
        m = self.mesh
        self.boundary("rho", m.nx, m.ny, NO_INVERT)
        self.boundary("e", m.nx, m.ny, NO_INVERT)

    def array_ptrs(self):
        return (ctypes.c_void_p * 16)(*[self.ptrs[k] for k in ARRAY_NAMES])

    def solve(self, tt):
        lib().solve_hydro_2d(self.mesh.ptr, tt,
                             *[ctypes.c_void_p(self.ptrs[k]) for k in ARRAY_NAMES],
                             None)

    def solve_static(self, tt, stream=None):
        """One step through flowhip_solve_hydro_2d_static (no allocation).
        Returns (status, mass, energy)."""
        mass, energy = ctypes.c_double(), ctypes.c_double()
        rc = lib().flowhip_solve_hydro_2d_static(
            self.mesh.ptr, tt, *[ctypes.c_void_p(self.ptrs[k]) for k in ARRAY_NAMES],
            ctypes.byref(mass), ctypes.byref(energy), stream)
        return rc, mass.value, energy.value

    def set_timestep(self, first_step):
        s = self.mesh.s
        lib().set_timestep(s.local_nx, s.local_ny,
                           ctypes.c_void_p(self.ptrs["Qxx"]),
                           ctypes.c_void_p(self.ptrs["Qyy"]),
                           ctypes.c_void_p(self.ptrs["rho"]),
                           ctypes.c_void_p(self.ptrs["e"]), self.mesh.ptr, None,
                           first_step, ctypes.cast(s.celldx, ctypes.c_void_p),
                           ctypes.cast(s.celldy, ctypes.c_void_p))

    def run(self, niters, tt0=0, pre_timestep=True):
        """main.c:64-115 through flowhip_run. Returns (steps, dt_pairs,
        mass, energy, wall_seconds); self.undefined then holds the undefined
        CFL candidates of the pre-loop call and of each step."""
        dt = np.zeros(2 * (niters + 1))
        mass = np.zeros(max(niters, 1))
        energy = np.zeros(max(niters, 1))
        und = np.zeros(niters + 1, dtype=np.int32)
        wall = ctypes.c_double()
        steps = lib().flowhip_run(self.mesh.ptr, self.array_ptrs(), niters, tt0,
                                  1 if pre_timestep else 0, _ptr(dt),
                                  _ptr(mass), _ptr(energy), ctypes.byref(wall),
                                  und.ctypes.data_as(ctypes.c_void_p))
        self.undefined = und[: steps + 1].copy()
        return (steps, dt[: 2 * (steps + 1)].reshape(-1, 2), mass[:steps],
                energy[:steps], wall.value)


class DeviceState32:
    """The FP32-state precision form: the same 16 arrays as FP32 device
    arrays of (L+1)^2 floats (flowhip_f32_alloc)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.n = (mesh.nx + 1) * (mesh.ny + 1)
        self.ptrs = {}
        for k in ARRAY_NAMES:
            p = ctypes.c_void_p()
            lib().flowhip_f32_alloc(ctypes.byref(p), self.n)
            self.ptrs[k] = p.value

    def close(self):
        for k, p in list(self.ptrs.items()):
            lib().flowhip_f32_free(ctypes.c_void_p(p))
        self.ptrs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, name, host):
        """Round host values (any float type) to FP32 and copy them in."""
        host = np.ascontiguousarray(host, dtype=np.float64).ravel()
        buf = np.zeros(self.n, dtype=np.float32)
        buf[: host.size] = host.astype(np.float32)
        lib().flowhip_f32_copy(ctypes.c_void_p(self.ptrs[name]),
                               buf.ctypes.data_as(ctypes.c_void_p), self.n, SEND)

    def zero(self, name):
        lib().flowhip_f32_zero(ctypes.c_void_p(self.ptrs[name]), self.n)

    def download(self, name):
        """The array widened to float64."""
        buf = np.empty(self.n, dtype=np.float32)
        lib().flowhip_f32_copy(ctypes.c_void_p(self.ptrs[name]),
                               buf.ctypes.data_as(ctypes.c_void_p), self.n, RECV)
        return buf.astype(np.float64)

    def download_all(self):
        return {k: self.download(k) for k in ARRAY_NAMES}

    def reflect_initial(self):
        m = self.mesh
        for k in ("rho", "e"):
            lib().flowhip_f32_boundary(m.ptr, ctypes.c_void_p(self.ptrs[k]),
                                       m.nx, m.ny, NO_INVERT)

    def array_ptrs(self):
        return (ctypes.c_void_p * 16)(*[self.ptrs[k] for k in ARRAY_NAMES])

    def run(self, niters, tt0=0, pre_timestep=True):
        """main.c:64-115 on FP32 state (flowhip_run_f32). Returns (steps,
        dt_pairs, mass, energy, wall_seconds); scalars are FP64."""
        dt = np.zeros(2 * (niters + 1))
        mass = np.zeros(max(niters, 1))
        energy = np.zeros(max(niters, 1))
        und = np.zeros(niters + 1, dtype=np.int32)
        wall = ctypes.c_double()
        steps = lib().flowhip_run_f32(self.mesh.ptr, self.array_ptrs(), niters,
                                      tt0, 1 if pre_timestep else 0, _ptr(dt),
                                      _ptr(mass), _ptr(energy), ctypes.byref(wall),
                                      und.ctypes.data_as(ctypes.c_void_p))
        self.undefined = und[: steps + 1].copy()
        return (steps, dt[: 2 * (steps + 1)].reshape(-1, 2), mass[:steps],
                energy[:steps], wall.value)


class Ensemble:
    """The batching form (flowhip_ensemble_*): independent meshes of uneven
    sizes advanced together, FP64 or FP32 state."""

    def __init__(self, sizes, width=1.0, max_dt=1.0e30, dt=1.0e-4, fp32=False):
        n = len(sizes)
        nx = (ctypes.c_int * n)(*[int(a) for a in sizes])
        ny = (ctypes.c_int * n)(*[int(a) for a in sizes])
        arr = lambda v: (ctypes.c_double * n)(*([float(v)] * n))
        self.ptr = lib().flowhip_ensemble_new(n, nx, ny, arr(width), arr(width),
                                              arr(max_dt), arr(dt), 1 if fp32 else 0)
        if not self.ptr:
            raise RuntimeError("flowhip_ensemble_new failed")
        self.n = n
        self.sizes = list(sizes)
        self.fp32 = fp32

    def close(self):
        if self.ptr:
            lib().flowhip_ensemble_delete(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def local_size(self, i):
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().flowhip_ensemble_local_size(self.ptr, i, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_state(self, i, rho, e):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        nx, ny = self.local_size(i)
        assert rho.size == e.size == (nx + 1) * (ny + 1)
        if lib().flowhip_ensemble_set_state(self.ptr, i, _ptr(rho), _ptr(e)) != 0:
            raise ValueError("bad instance")

    def get(self, i, name):
        nx, ny = self.local_size(i)
        buf = np.empty((nx + 1) * (ny + 1))
        lib().flowhip_ensemble_get_array(self.ptr, i, ARRAY_NAMES.index(name), _ptr(buf))
        return buf

    def get_all(self, i):
        return {k: self.get(i, k) for k in ARRAY_NAMES}

    def set_graphs(self, on):
        lib().flowhip_ensemble_set_graphs(self.ptr, 1 if on else 0)

    def set_schedule(self, chunk_rows, order, fused_momentum=True):
        if lib().flowhip_ensemble_set_schedule(self.ptr, chunk_rows, order,
                                               1 if fused_momentum else 0) != 0:
            raise ValueError("bad schedule")

    def schedule(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().flowhip_ensemble_get_schedule(self.ptr, ctypes.byref(a), ctypes.byref(b),
                                            ctypes.byref(c))
        return a.value, b.value, bool(c.value)

    def autotune(self, trial_steps=3, tt0=0):
        ms = ctypes.c_double()
        lib().flowhip_ensemble_autotune(self.ptr, trial_steps, tt0, ctypes.byref(ms))
        return self.schedule(), ms.value

    def run(self, niters, tt0=0, pre_timestep=True):
        """Returns (dt [inst][niters+1][2], mass [inst][niters],
        energy [inst][niters], wall_seconds)."""
        dt = np.zeros(self.n * 2 * (niters + 1))
        mass = np.zeros(max(self.n * niters, 1))
        energy = np.zeros(max(self.n * niters, 1))
        und = np.zeros(max(self.n * niters, 1), dtype=np.int32)
        wall = ctypes.c_double()
        lib().flowhip_ensemble_run(self.ptr, niters, tt0, 1 if pre_timestep else 0,
                                   _ptr(dt), _ptr(mass), _ptr(energy), ctypes.byref(wall),
                                   und.ctypes.data_as(ctypes.c_void_p))
        self.undefined = und[: self.n * niters].reshape(self.n, niters).copy()
        return (dt.reshape(self.n, niters + 1, 2),
                mass[: self.n * niters].reshape(self.n, niters),
                energy[: self.n * niters].reshape(self.n, niters), wall.value)


def device_name():
    buf = ctypes.create_string_buffer(256)
    lib().flowhip_device_name(buf, 256)
    return buf.value.decode()


def sync():
    lib().flowhip_device_sync()
