"""ctypes binding of libsphgpu.so (the C-ABI declared in include/sphgpu.h).

The product path is the HIP library; this module only marshals numpy arrays
across the C-ABI.  Loading fails loudly when the library is missing, and the
compute entry points return SPHGPU_E_DEVICE when no gfx9 device is present:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsphgpu.so"

WCSPH, DFSPH, IISPH, PCISPH, IISPH_DF = 0, 1, 2, 3, 4
WARM_NONE, WARM_PREVIOUS, WARM_LEARNED = 0, 1, 2
WIN_DENSITY, WIN_DIVERGENCE, WIN_VELOCITY, WIN_SUMDIJ, WIN_PRESSURE = 1, 2, 4, 8, 16
F_POS, F_VEL, F_RHO, F_PRESSURE, F_ACC, F_ALPHA, F_KAPPA_SUM, F_PHASE, F_KAPPAV_SUM = range(9)
FIELD_COMPONENTS = {F_POS: 3, F_VEL: 3, F_RHO: 1, F_PRESSURE: 1, F_ACC: 3, F_ALPHA: 1,
                    F_KAPPA_SUM: 1, F_PHASE: 1, F_KAPPAV_SUM: 1}

ERRORS = {0: "OK", -1: "E_ARG", -2: "E_STATE", -3: "E_DEVICE", -4: "E_CAPACITY", -5: "E_DIVERGED",
          -6: "E_CALLBACK"}

POS_POSITION, POS_VELOCITY, POS_DENSITY, POS_PRESURE, POS_ACCELERATION = 1, 2, 8, 16, 32

# int (*)(sphgpu_handle*, void* user)
UPDATE_SOLIDS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)


class Params(C.Structure):
    _fields_ = [
        ("solver", C.c_int),
        ("spacing_r", C.c_float),
        ("smooth_radius_h", C.c_float),
        ("space_min", C.c_float * 3),
        ("space_max", C.c_float * 3),
        ("dt", C.c_float),
        ("gravity", C.c_float * 3),
        ("sound_speed", C.c_float),
        ("gamma", C.c_float),
        ("nu", C.c_float),
        ("cfl", C.c_float),
        ("dt_min", C.c_float),
        ("density_tol", C.c_float),
        ("divergence_tol", C.c_float),
        ("min_iter_density", C.c_int),
        ("max_iter_density", C.c_int),
        ("min_iter_div", C.c_int),
        ("max_iter_div", C.c_int),
        ("omega", C.c_float),
        ("warm_start", C.c_int),
        ("num_scenes", C.c_int),
        ("learn_steps", C.c_int),
        ("learn_iters", C.c_int),
        ("learn_batch", C.c_int),
        ("learn_rate", C.c_float),
        ("learn_over_weight", C.c_float),
        ("warm_scale", C.c_float),
        ("use_graph", C.c_int),
        ("boundary_density", C.c_float),
        ("adjust_density", C.c_int),
        ("fluid_capacity", C.c_int),
        ("learn_scale", C.c_float),
        ("multiphase_consistent", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iter_density", C.c_int),
        ("iter_divergence", C.c_int),
        ("density_err", C.c_double),
        ("divergence_err", C.c_double),
        ("rho_mean_err", C.c_double),
        ("rho_max_err", C.c_double),
        ("kinetic_energy", C.c_double),
        ("dt", C.c_float),
        ("zero_dis", C.c_int),
        ("too_many_neigb", C.c_int),
        ("removed", C.c_int),
        ("max_iter_hit", C.c_int),
        ("nan_count", C.c_int),
        ("num_fluid", C.c_longlong),
        ("list_overflow", C.c_int),
        ("reserved", C.c_int * 5),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_LP = C.POINTER(C.c_longlong)
_DP = C.POINTER(C.c_double)

# exported symbol -> (restype, argtypes); every declaration of include/sphgpu.h
SIGNATURES = {
    "sphgpu_default_params": (None, [C.POINTER(Params)]),
    "sphgpu_abi_version": (C.c_int, []),
    "sphgpu_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "sphgpu_destroy": (None, [C.c_void_p]),
    "sphgpu_last_error": (C.c_char_p, [C.c_void_p]),
    "sphgpu_device_available": (C.c_int, []),
    "sphgpu_set_scene_domain": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP]),
    "sphgpu_add_phase": (C.c_int, [C.c_void_p, C.c_float, C.c_float, _IP]),
    "sphgpu_add_fluid": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_longlong, _FP, _FP, _LP]),
    "sphgpu_add_solid": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong, _FP, C.c_float, _IP]),
    "sphgpu_set_fluid_beta": (C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, _FP]),
    "sphgpu_append_fluid": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_longlong, _FP, _FP, _LP]),
    "sphgpu_add_rigid": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong, _FP, C.c_float, _IP]),
    "sphgpu_set_solid_pose": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP, _FP, _FP]),
    "sphgpu_get_solid_force_torque": (C.c_int, [C.c_void_p, C.c_int, _DP, _DP]),
    "sphgpu_set_update_solids": (C.c_int, [C.c_void_p, UPDATE_SOLIDS_FN, C.c_void_p]),
    "sphgpu_init": (C.c_int, [C.c_void_p]),
    "sphgpu_step": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "sphgpu_scene_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Stats)]),
    "sphgpu_get_step_dt": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    "sphgpu_snapshot_begin": (C.c_int, [C.c_void_p]),
    "sphgpu_snapshot_wait": (C.c_int, [C.c_void_p]),
    "sphgpu_write_pos": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_uint, C.c_char_p]),
    "sphgpu_write_solid_pos": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p]),
    "sphgpu_write_json": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "sphgpu_read_pos": (C.c_int, [C.c_char_p, _FP, C.c_longlong, C.POINTER(C.c_longlong),
                                  C.POINTER(C.c_longlong), C.c_char_p, C.c_int]),
    "sphgpu_surfacer_create": (C.c_int, [_FP, _FP, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_void_p)]),
    "sphgpu_surfacer_destroy": (None, [C.c_void_p]),
    "sphgpu_surfacer_last_error": (C.c_char_p, [C.c_void_p]),
    "sphgpu_surfacer_grid": (C.c_int, [C.c_void_p, _DP, C.POINTER(C.c_int), _DP]),
    "sphgpu_surfacer_aniso": (C.c_int, [C.c_void_p, C.c_longlong, _FP]),
    "sphgpu_surfacer_isotro": (C.c_int, [C.c_void_p, C.c_longlong, _FP]),
    "sphgpu_surfacer_particles": (C.c_int, [C.c_void_p, _FP, _FP, _FP, C.POINTER(C.c_int)]),
    "sphgpu_surfacer_get_field": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_surfacer_set_field": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_surfacer_marching_cubes": (C.c_int, [C.c_void_p, C.c_float, _FP, C.c_longlong,
                                                 C.POINTER(C.c_longlong)]),
    "sphgpu_marching_cubes_table": (C.c_int, [C.POINTER(C.c_ushort), C.POINTER(C.c_ubyte),
                                              C.POINTER(C.c_byte)]),
    "sphgpu_synchronize": (C.c_int, [C.c_void_p]),
    "sphgpu_num_fluid": (C.c_longlong, [C.c_void_p]),
    "sphgpu_get_fluid": (C.c_int, [C.c_void_p, C.c_int, _FP]),
    "sphgpu_get_fluid_ids": (C.c_int, [C.c_void_p, _LP]),
    "sphgpu_get_boundary_volume": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_get_neighbours": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, _LP, _FP]),
    "sphgpu_profile_kernel": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sphgpu_kernel_timing": (C.c_int, [C.c_void_p, _LP, _DP, _DP]),
    "sphgpu_copy_probe": (C.c_int, [C.c_longlong, _DP]),
    "sphgpu_get_tuning": (C.c_int, [C.c_void_p, _IP, _FP]),
    "sphgpu_get_list_form": (C.c_int, [C.c_void_p, _IP, _IP, _IP]),
    "sphgpu_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "sphgpu_set_lds_windows": (C.c_int, [C.c_void_p, C.c_int]),
    "sphgpu_get_lds_windows": (C.c_int, [C.c_void_p, _IP, _FP, _IP]),
    "sphgpu_mlp_create": (C.c_int, [C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]),
    "sphgpu_mlp_destroy": (None, [C.c_void_p]),
    "sphgpu_mlp_set_params": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_mlp_get_params": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_mlp_forward": (C.c_int, [C.c_void_p, C.c_longlong, _FP, _FP]),
    "sphgpu_mlp_gradient": (C.c_int, [C.c_void_p, C.c_longlong, _FP, _FP, _FP, _DP]),
    "sphgpu_mlp_set_loss": (C.c_int, [C.c_void_p, C.c_float]),
    "sphgpu_mlp_gradient_device": (C.c_int, [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_float,
                                             C.c_void_p, C.c_void_p, C.c_void_p]),
    "sphgpu_mlp_train": (C.c_int, [C.c_void_p, C.c_longlong, _FP, _FP, C.c_int, C.c_float, _DP]),
    "sphgpu_get_mlp_params": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_set_mlp_params": (C.c_int, [C.c_void_p, _FP]),
    "sphgpu_scene_lattice": (C.c_longlong, [_IP, _FP, C.c_float, C.c_float, C.c_ulonglong,
                                            C.c_ulonglong, _FP]),
    "sphgpu_scene_box": (C.c_longlong, [_FP, _FP, C.c_float, C.c_int, _FP]),
}

_lib = None


def load() -> C.CDLL:
    """Loads the in-tree libsphgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(the HIP extension is required; there is no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_FP)


class SphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def default_params(**kw) -> Params:
    p = Params()
    load().sphgpu_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("space_min", "space_max", "gravity"):
            getattr(p, k)[:] = [float(t) for t in v]
        else:
            setattr(p, k, v)
    return p


def device_available() -> bool:
    return bool(load().sphgpu_device_available())


def copy_probe(nbytes: int) -> float:
    """Plain device-to-device copy of nbytes (counter calibration); returns ms."""
    ms = C.c_double()
    rc = load().sphgpu_copy_probe(int(nbytes), C.byref(ms))
    if rc:
        raise SphError(rc, "copy probe failed")
    return ms.value


def pos_record_floats(mask: str) -> int:
    """Floats per record named by a .pos component line (fUtility.h:53-63)."""
    size = 0
    for name, nbytes in (("position", 12), ("velocity", 12), ("color", 16), ("density", 4),
                         ("presure", 4), ("acceleration", 12), ("dt", 4)):
        if name in mask:
            size += nbytes
    return size // 4


def read_pos(path):
    """read_fluid_particles (fUtility.h:486-510) through the library: returns
    (frame, component line, records[n, floats per record])."""
    lib = load()
    n, fr = C.c_longlong(), C.c_longlong()
    mask = C.create_string_buffer(512)
    rc = lib.sphgpu_read_pos(str(path).encode(), None, 0, C.byref(n), C.byref(fr), mask, 512)
    if rc != 0:
        raise SphError(rc, f"read_pos {path}")
    per = pos_record_floats(mask.value.decode())
    data = np.zeros(n.value * per, np.float32)
    rc = lib.sphgpu_read_pos(str(path).encode(), _fp(data), data.size, C.byref(n), C.byref(fr), mask, 512)
    if rc != 0:
        raise SphError(rc, f"read_pos {path}")
    return fr.value, mask.value.decode(), data.reshape(n.value, per)


def marching_cubes_table():
    """(edges[256] uint16, ntris[256] uint8, tris[256, 16] int8) of the surfacer."""
    lib = load()
    e = np.zeros(256, np.uint16)
    n = np.zeros(256, np.uint8)
    t = np.zeros((256, 16), np.int8)
    lib.sphgpu_marching_cubes_table(e.ctypes.data_as(C.POINTER(C.c_ushort)), n.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                    t.ctypes.data_as(C.POINTER(C.c_byte)))
    return e, n, t


class Surfacer:
    """The reference's offline surfacer (surface_field + marchingCubes) on the device."""

    def __init__(self, mn, mx, radius_p, radius_h, cell_size):
        self.lib = load()
        h = C.c_void_p()
        a, b = _f32(mn), _f32(mx)
        rc = self.lib.sphgpu_surfacer_create(_fp(a), _fp(b), radius_p, radius_h, cell_size, C.byref(h))
        if rc != 0:
            raise SphError(rc, "surfacer_create")
        self.h = h.value
        gmin = np.zeros(3)
        gnum = (C.c_int * 3)()
        cell = C.c_double()
        self.lib.sphgpu_surfacer_grid(self.h, gmin.ctypes.data_as(_DP), gnum, C.byref(cell))
        self.grid_min, self.grid_num, self.cell = gmin, tuple(gnum), cell.value

    def _check(self, rc, what):
        if rc != 0:
            raise SphError(rc, f"{what}: {self.lib.sphgpu_surfacer_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.sphgpu_surfacer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def aniso(self, xyz):
        x = _f32(xyz).reshape(-1, 3)
        self._check(self.lib.sphgpu_surfacer_aniso(self.h, len(x), _fp(x)), "aniso")

    def isotro(self, xyz):
        x = _f32(xyz).reshape(-1, 3)
        self._check(self.lib.sphgpu_surfacer_isotro(self.h, len(x), _fp(x)), "isotro")

    def particles(self, n):
        xb, G, det = np.zeros((n, 3), np.float32), np.zeros((n, 3, 3), np.float32), np.zeros(n, np.float32)
        nb = np.zeros(n, np.int32)
        self._check(self.lib.sphgpu_surfacer_particles(self.h, _fp(xb), _fp(G), _fp(det),
                                                       nb.ctypes.data_as(C.POINTER(C.c_int))), "particles")
        return xb, G, det, nb

    def field(self) -> np.ndarray:
        nx, ny, nz = self.grid_num
        f = np.zeros(nx * ny * nz, np.float32)
        self._check(self.lib.sphgpu_surfacer_get_field(self.h, _fp(f)), "get_field")
        return f.reshape(nz, ny, nx)

    def set_field(self, f):
        a = _f32(f).ravel()
        assert a.size == int(np.prod(self.grid_num))
        self._check(self.lib.sphgpu_surfacer_set_field(self.h, _fp(a)), "set_field")

    def marching_cubes(self, iso=0.5) -> np.ndarray:
        n = C.c_longlong()
        self._check(self.lib.sphgpu_surfacer_marching_cubes(self.h, iso, None, 0, C.byref(n)), "marching_cubes")
        out = np.zeros((max(n.value, 1), 3, 3), np.float32)
        self._check(self.lib.sphgpu_surfacer_marching_cubes(self.h, iso, _fp(out), n.value, C.byref(n)),
                    "marching_cubes")
        return out[:n.value]


def lattice(n, origin, spacing, jitter=0.0, seed=0, stream=0) -> np.ndarray:
    lib = load()
    n3 = (C.c_int * 3)(*[int(t) for t in n])
    o3 = (C.c_float * 3)(*[float(t) for t in origin])
    total = int(n[0]) * int(n[1]) * int(n[2])
    out = np.empty((total, 3), np.float32)
    got = lib.sphgpu_scene_lattice(n3, o3, float(spacing), float(jitter), int(seed), int(stream), _fp(out))
    assert got == total
    return out


def box_walls(center, half, spacing, have_top=True) -> np.ndarray:
    lib = load()
    c3 = (C.c_float * 3)(*[float(t) for t in center])
    h3 = (C.c_float * 3)(*[float(t) for t in half])
    n = lib.sphgpu_scene_box(c3, h3, float(spacing), int(bool(have_top)), None)
    out = np.empty((n, 3), np.float32)
    lib.sphgpu_scene_box(c3, h3, float(spacing), int(bool(have_top)), _fp(out))
    return out


MLP_PARAMS = 8961


class Mlp:
    """Standalone per-particle MLP of the C-ABI (8 -> 64 -> 64 -> 64 -> 1)."""

    def __init__(self, capacity: int, seed: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.sphgpu_mlp_create(int(capacity), int(seed), C.byref(h))
        if rc != 0:
            raise SphError(rc, "sphgpu_mlp_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.sphgpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise SphError(rc, what)

    def set_params(self, p):
        a = _f32(p)
        assert a.size == MLP_PARAMS
        self._check(self.lib.sphgpu_mlp_set_params(self.h, _fp(a)), "mlp_set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        self._check(self.lib.sphgpu_mlp_get_params(self.h, _fp(out)), "mlp_get_params")
        return out

    def forward(self, feat) -> np.ndarray:
        f = _f32(feat).reshape(-1, 8)
        out = np.empty(len(f), np.float32)
        self._check(self.lib.sphgpu_mlp_forward(self.h, len(f), _fp(f), _fp(out)), "mlp_forward")
        return out

    def gradient(self, feat, target):
        f = _f32(feat).reshape(-1, 8)
        t = _f32(target)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_double()
        self._check(self.lib.sphgpu_mlp_gradient(self.h, len(f), _fp(f), _fp(t), _fp(g), C.byref(loss)),
                    "mlp_gradient")
        return g, loss.value

    def set_loss(self, over_weight: float):
        self._check(self.lib.sphgpu_mlp_set_loss(self.h, float(over_weight)), "mlp_set_loss")

    def train(self, feat, target, iters, lr):
        f = _f32(feat).reshape(-1, 8)
        t = _f32(target)
        loss = C.c_double()
        self._check(self.lib.sphgpu_mlp_train(self.h, len(f), _fp(f), _fp(t), int(iters), float(lr),
                                              C.byref(loss)), "mlp_train")
        return loss.value


class Simulation:
    """One handle of the C-ABI (a SphBase-derived solver object)."""

    def __init__(self, params: Params):
        self.lib = load()
        self.params = params
        h = C.c_void_p()
        rc = self.lib.sphgpu_create(C.byref(params), C.byref(h))
        if rc != 0:
            raise SphError(rc, "sphgpu_create rejected the parameters")
        self.h = h

    # -- helpers
    def _check(self, rc: int, what: str):
        if rc != 0:
            err = SphError(rc, f"{what}: {self.lib.sphgpu_last_error(self.h).decode()}")
            cause, self._cb_error = getattr(self, "_cb_error", None), None
            if cause is not None:
                raise err from cause
            raise err

    def close(self):
        if self.h:
            self.lib.sphgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scene
    def set_scene_domain(self, scene, mn, mx):
        a, b = _f32(mn), _f32(mx)
        self._check(self.lib.sphgpu_set_scene_domain(self.h, scene, _fp(a), _fp(b)), "set_scene_domain")

    def add_phase(self, rest_density, visc_alpha) -> int:
        ph = C.c_int()
        self._check(self.lib.sphgpu_add_phase(self.h, rest_density, visc_alpha, C.byref(ph)), "add_phase")
        return ph.value

    def add_fluid(self, xyz, vel=None, phase=0, scene=0) -> int:
        x = _f32(xyz).reshape(-1, 3)
        v = None if vel is None else _f32(vel).reshape(-1, 3)
        first = C.c_longlong()
        self._check(self.lib.sphgpu_add_fluid(self.h, scene, phase, len(x), _fp(x),
                                              None if v is None else _fp(v), C.byref(first)),
                    "add_fluid")
        return first.value

    def append_fluid(self, xyz, vel=None, phase=0, scene=0) -> int:
        """Fluid particles added between steps (an emitter); returns the first new id."""
        x = _f32(xyz).reshape(-1, 3)
        v = None if vel is None else _f32(vel).reshape(-1, 3)
        first = C.c_longlong()
        self._check(self.lib.sphgpu_append_fluid(self.h, scene, phase, len(x), _fp(x),
                                                 None if v is None else _fp(v), C.byref(first)), "append_fluid")
        return first.value

    def set_fluid_beta(self, first, beta):
        """Rest-density factor beta of the fluid particles [first, first + len(beta))
        in add order (IISph DF_* solvers); before init."""
        b = _f32(beta).ravel()
        self._check(self.lib.sphgpu_set_fluid_beta(self.h, int(first), len(b), _fp(b)), "set_fluid_beta")

    def add_solid(self, xyz, visc_alpha=0.0, scene=0) -> int:
        x = _f32(xyz).reshape(-1, 3)
        sid = C.c_int()
        self._check(self.lib.sphgpu_add_solid(self.h, scene, len(x), _fp(x), visc_alpha, C.byref(sid)),
                    "add_solid")
        return sid.value

    def add_rigid(self, local_xyz, visc_alpha=0.0, scene=0) -> int:
        x = _f32(local_xyz).reshape(-1, 3)
        sid = C.c_int()
        self._check(self.lib.sphgpu_add_rigid(self.h, scene, len(x), _fp(x), visc_alpha, C.byref(sid)),
                    "add_rigid")
        return sid.value

    def set_solid_pose(self, solid, T16, com, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)):
        t, c, vv, ww = (_f32(a).ravel() for a in (T16, com, v, w))
        self._check(self.lib.sphgpu_set_solid_pose(self.h, solid, _fp(t), _fp(c), _fp(vv), _fp(ww)),
                    "set_solid_pose")

    def solid_force_torque(self, solid):
        f = np.zeros(3)
        t = np.zeros(3)
        self._check(self.lib.sphgpu_get_solid_force_torque(self.h, solid, f.ctypes.data_as(_DP),
                                                           t.ctypes.data_as(_DP)), "solid_force_torque")
        return f, t

    def set_update_solids(self, fn):
        """fn(sim) runs inside step() where the reference calls
        SphWithBullet::updateSolids; it may read solid_force_torque and call
        set_solid_pose.  An exception in fn fails the step (E_CALLBACK) and
        is re-raised from step().  fn=None removes it."""
        if fn is None:
            self._update_cb = None
            self._check(self.lib.sphgpu_set_update_solids(self.h, UPDATE_SOLIDS_FN(), None), "set_update_solids")
            return

        def trampoline(_h, _user):
            try:
                fn(self)
                return 0
            except BaseException as e:        # no exception may cross the C frame
                self._cb_error = e
                return 1

        self._update_cb = UPDATE_SOLIDS_FN(trampoline)   # kept alive with the handle
        self._check(self.lib.sphgpu_set_update_solids(self.h, self._update_cb, None), "set_update_solids")

    def init(self):
        self._check(self.lib.sphgpu_init(self.h), "init")

    # -- stepping
    def step(self, stats: bool = True):
        """One step; returns the statistics dict, or None (and does not wait
        for the device) when stats is False."""
        if not stats:
            self._check(self.lib.sphgpu_step(self.h, None), "step")
            return None
        st = Stats()
        self._check(self.lib.sphgpu_step(self.h, C.byref(st)), "step")
        return st.as_dict()

    # -- frame output (the reference's .pos / .json formats)
    def snapshot_begin(self):
        self._check(self.lib.sphgpu_snapshot_begin(self.h), "snapshot_begin")

    def snapshot_wait(self):
        self._check(self.lib.sphgpu_snapshot_wait(self.h), "snapshot_wait")

    def write_pos(self, path, phase=0, components=None, log_name=""):
        comps = POS_POSITION if components is None else components
        self._check(self.lib.sphgpu_write_pos(self.h, str(path).encode(), phase, comps, log_name.encode()),
                    "write_pos")

    def write_solid_pos(self, path, solid, log_name=""):
        self._check(self.lib.sphgpu_write_solid_pos(self.h, str(path).encode(), solid, log_name.encode()),
                    "write_solid_pos")

    def write_json(self, path, phase=0):
        self._check(self.lib.sphgpu_write_json(self.h, str(path).encode(), phase), "write_json")

    def step_dt(self, scene=0) -> float:
        """dt of the current step (inside the update-solids callback) or the last one."""
        dt = C.c_float()
        self._check(self.lib.sphgpu_get_step_dt(self.h, scene, C.byref(dt)), "step_dt")
        return dt.value

    def scene_stats(self, scene) -> dict:
        st = Stats()
        self._check(self.lib.sphgpu_scene_stats(self.h, scene, C.byref(st)), "scene_stats")
        return st.as_dict()

    def synchronize(self):
        self._check(self.lib.sphgpu_synchronize(self.h), "synchronize")

    # -- readback
    def num_fluid(self) -> int:
        return int(self.lib.sphgpu_num_fluid(self.h))

    def get(self, field) -> np.ndarray:
        n = self.num_fluid()
        comp = FIELD_COMPONENTS[field]
        out = np.empty((n, comp), np.float32)
        self._check(self.lib.sphgpu_get_fluid(self.h, field, _fp(out)), "get_fluid")
        return out if comp > 1 else out[:, 0]

    def ids(self) -> np.ndarray:
        out = np.empty(self.num_fluid(), np.int64)
        self._check(self.lib.sphgpu_get_fluid_ids(self.h, out.ctypes.data_as(_LP)), "get_fluid_ids")
        return out

    def boundary_volume(self, nb) -> np.ndarray:
        out = np.empty(nb, np.float32)
        self._check(self.lib.sphgpu_get_boundary_volume(self.h, _fp(out)), "get_boundary_volume")
        return out

    def neighbours(self, cap=96):
        n = self.num_fluid()
        counts = np.zeros(n, np.int32)
        kinds = np.zeros((n, cap), np.int32)
        ids = np.zeros((n, cap), np.int64)
        dis = np.zeros((n, cap), np.float32)
        self._check(self.lib.sphgpu_get_neighbours(
            self.h, cap, counts.ctypes.data_as(_IP), kinds.ctypes.data_as(_IP),
            ids.ctypes.data_as(_LP), _fp(dis)), "get_neighbours")
        return counts, kinds, ids, dis

    # -- learned warm start
    def mlp_params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        self._check(self.lib.sphgpu_get_mlp_params(self.h, _fp(out)), "get_mlp_params")
        return out

    def set_mlp_params(self, p):
        a = _f32(p)
        self._check(self.lib.sphgpu_set_mlp_params(self.h, _fp(a)), "set_mlp_params")

    # -- measurement
    def tuning(self) -> dict:
        ch = (C.c_int * 2)()
        ms = (C.c_float * 4)()
        self._check(self.lib.sphgpu_get_tuning(self.h, ch, ms), "get_tuning")
        return {"tile_map_xcd": bool(ch[0]), "list_stage_groups": ch[1],
                "ms": {"density_round_robin": ms[0], "density_xcd": ms[1],
                       "lists_6": ms[2], "lists_8": ms[3]}}

    def set_tuning(self, tile_map=None, stage_groups=None, stream_lists=None):
        """Fixes autotuner choices (None: timed at the first step)."""
        f = lambda v: -1 if v is None else int(v)
        self._check(self.lib.sphgpu_set_tuning(self.h, f(tile_map), f(stage_groups), f(stream_lists)),
                    "set_tuning")

    def set_lds_windows(self, on: bool):
        """Every solver-loop pass reads packed tiles' neighbours from LDS
        windows (on) or gathers them per lane (off), instead of the
        autotuner's per-pass choice; the results are bit-identical."""
        self._check(self.lib.sphgpu_set_lds_windows(self.h, int(bool(on))), "set_lds_windows")

    def lds_windows(self) -> dict:
        """Passes with their LDS window on (WIN_* bits) and the autotuner's
        timings per pass: (gathers ms, window ms), None where not timed."""
        mask, ms, bl = C.c_int(), (C.c_float * 10)(), (C.c_int * 5)()
        self._check(self.lib.sphgpu_get_lds_windows(self.h, C.byref(mask), ms, bl), "get_lds_windows")
        names = ("density", "divergence", "velocity", "sumdij", "pressure")
        return {"mask": mask.value,
                "ms": {n: (ms[2 * k], ms[2 * k + 1]) if ms[2 * k] > 0 else None for k, n in enumerate(names)},
                "blocks_per_cu": {n: bl[k] for k, n in enumerate(names)}}

    def list_form(self) -> dict:
        """Tiles holding particles, how many of them have packed (16-bit) lists,
        and whether the passes stream the lists (autotuned)."""
        p, t, st = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.sphgpu_get_list_form(self.h, C.byref(p), C.byref(t), C.byref(st)), "get_list_form")
        return {"packed_tiles": p.value, "tiles": t.value, "streamed": bool(st.value)}

    def profile_kernel(self, name):
        self._check(self.lib.sphgpu_profile_kernel(self.h, None if name is None else name.encode()),
                    "profile_kernel")

    def kernel_timing(self):
        n, ms, b = C.c_longlong(), C.c_double(), C.c_double()
        self._check(self.lib.sphgpu_kernel_timing(self.h, C.byref(n), C.byref(ms), C.byref(b)),
                    "kernel_timing")
        return n.value, ms.value, b.value
