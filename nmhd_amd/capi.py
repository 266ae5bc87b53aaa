"""ctypes binding of libnmhd.so (C ABI in include/nmhd.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback on the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnmhd.so")

NMHD_OK = 0
ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "EIO", 5: "ESTATE"}
PATH_AUTO, PATH_UNFUSED, PATH_TILED, PATH_STRICT = 0, 1, 2, 3
RK2, EULER1 = 0, 1

# Every entry point declared in include/nmhd.h.
EXPORTS = (
    "nmhd_create", "nmhd_destroy", "nmhd_last_error", "nmhd_get_geometry",
    "nmhd_set_state", "nmhd_get_state", "nmhd_set_state_device",
    "nmhd_get_state_device", "nmhd_set_time", "nmhd_get_time",
    "nmhd_fill_ghosts", "nmhd_relative_displacement", "nmhd_cfl_dt", "nmhd_rhs",
    "nmhd_step", "nmhd_run", "nmhd_reset_run", "nmhd_enqueue_steps",
    "nmhd_synchronize", "nmhd_hlld_batch", "nmhd_timing_enable",
    "nmhd_timing_reset", "nmhd_timing_get", "nmhd_set_path", "nmhd_get_path",
    "nmhd_set_tiles", "nmhd_get_tiles", "nmhd_get_diagnostics", "nmhd_batch_create", "nmhd_batch_destroy", "nmhd_batch_last_error",
    "nmhd_batch_get_geometry", "nmhd_batch_set_state", "nmhd_batch_get_state",
    "nmhd_batch_set_time", "nmhd_batch_get_time", "nmhd_batch_get_status",
    "nmhd_batch_set_tiles", "nmhd_batch_set_launch", "nmhd_batch_get_launch",
    "nmhd_batch_autotune", "nmhd_batch_run",
    "nmhd_adj_create", "nmhd_adj_destroy", "nmhd_adj_last_error", "nmhd_adj_get_geometry",
    "nmhd_adj_set_stream", "nmhd_adj_synchronize", "nmhd_adj_cfl_dt", "nmhd_adj_forward",
    "nmhd_adj_reverse", "nmhd_adj_forward_stages", "nmhd_adj_reverse_stages",
    "nmhd_adj_forward_noalloc",
)
TILE_64x4, TILE_32x8, TILE_16x16 = 0, 1, 2
# enum nmhd_fix: corrections of reference quirks (0 = the reference's behaviour)
FIX_CFL_SPEED, FIX_SHEAR_INTERP, FIX_FACE_POSITION, FIX_ALL = 1, 2, 4, 7


class NmhdDesc(ctypes.Structure):
    _fields_ = [
        ("global_size", ctypes.c_int * 3),
        ("margin", ctypes.c_int * 3),
        ("global_extent", ctypes.c_float * 3),
        ("gamma", ctypes.c_float),
        ("cfl_factor", ctypes.c_float),
        ("source", ctypes.c_int),
        ("omega", ctypes.c_float),
        ("qshear", ctypes.c_float),
        ("shear_y", ctypes.c_int),
        ("recon", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("path", ctypes.c_int),
        ("fixes", ctypes.c_int),
    ]


class NmhdDiagnostics(ctypes.Structure):
    _fields_ = [
        ("b2_mean", ctypes.c_double * 3),
        ("total", ctypes.c_double * 8),
        ("divb_max", ctypes.c_double),
    ]


class NmhdGeometry(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int * 3),
        ("ncell", ctypes.c_size_t),
        ("extent", ctypes.c_float * 3),
        ("local_extent", ctypes.c_float * 3),
        ("d", ctypes.c_float * 3),
    ]


_FP = ctypes.POINTER(ctypes.c_float)
_FPP = ctypes.POINTER(_FP)
_VP = ctypes.c_void_p
_IP = ctypes.POINTER(ctypes.c_int)
_lib = None


class NmhdError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NmhdError("libnmhd.so is not built (run __graft_entry__.build() or "
                        "python nmhd_amd/build.py); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "nmhd_create": [ctypes.POINTER(NmhdDesc), ctypes.POINTER(_VP)],
        "nmhd_destroy": [_VP],
        "nmhd_get_geometry": [_VP, ctypes.POINTER(NmhdGeometry)],
        "nmhd_set_state": [_VP, _FPP],
        "nmhd_get_state": [_VP, _FPP],
        "nmhd_set_state_device": [_VP, _FPP],
        "nmhd_get_state_device": [_VP, _FPP],
        "nmhd_set_time": [_VP, ctypes.c_double, ctypes.c_int],
        "nmhd_get_time": [_VP, ctypes.POINTER(ctypes.c_double), _IP],
        "nmhd_fill_ghosts": [_VP],
        "nmhd_relative_displacement": [_VP, ctypes.c_double, _FP, _IP, _FP],
        "nmhd_cfl_dt": [_VP, ctypes.c_float, _FP],
        "nmhd_rhs": [_VP, ctypes.c_float, _FPP],
        "nmhd_step": [_VP, ctypes.c_float, _FP, _IP],
        "nmhd_run": [_VP, ctypes.c_float, ctypes.c_int, _IP, _IP],
        "nmhd_reset_run": [_VP],
        "nmhd_enqueue_steps": [_VP, ctypes.c_float, ctypes.c_int],
        "nmhd_synchronize": [_VP, _IP, _IP],
        "nmhd_hlld_batch": [ctypes.c_size_t, ctypes.c_float, ctypes.c_int, _FPP, _FPP],
        "nmhd_timing_enable": [_VP, ctypes.c_int],
        "nmhd_timing_reset": [_VP],
        "nmhd_timing_get": [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)],
        "nmhd_set_path": [_VP, ctypes.c_int],
        "nmhd_get_path": [_VP, _IP],
        "nmhd_set_tiles": [_VP, ctypes.c_int, ctypes.c_int],
        "nmhd_get_diagnostics": [_VP, ctypes.POINTER(NmhdDiagnostics)],
        "nmhd_get_tiles": [_VP, _IP, _IP],
        "nmhd_batch_create": [ctypes.POINTER(NmhdDesc), ctypes.c_int, ctypes.POINTER(_VP)],
        "nmhd_batch_destroy": [_VP],
        "nmhd_batch_get_geometry": [_VP, ctypes.c_int, ctypes.POINTER(NmhdGeometry)],
        "nmhd_batch_set_state": [_VP, ctypes.c_int, _FPP],
        "nmhd_batch_get_state": [_VP, ctypes.c_int, _FPP],
        "nmhd_batch_set_time": [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_int],
        "nmhd_batch_get_time": [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _IP],
        "nmhd_batch_get_status": [_VP, ctypes.c_int, _IP, _IP, _FP],
        "nmhd_batch_set_tiles": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "nmhd_batch_set_launch": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "nmhd_batch_get_launch": [_VP, _IP, _IP, _IP, _IP],
        "nmhd_batch_autotune": [_VP, ctypes.c_char_p, _IP],
        "nmhd_batch_run": [_VP, ctypes.c_float, ctypes.c_int, ctypes.c_int, _IP, _IP],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.nmhd_last_error.argtypes = [_VP]
    L.nmhd_last_error.restype = ctypes.c_char_p
    L.nmhd_batch_last_error.argtypes = [_VP]
    L.nmhd_batch_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _ptrs(arrays):
    arr = (_FP * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = ctypes.cast(ctypes.c_void_p(a), _FP) if isinstance(a, int) else \
            a.ctypes.data_as(_FP)
    return arr


def _check(rc, handle=None, batch=False):
    if rc != NMHD_OK:
        L = load()
        msg = (L.nmhd_batch_last_error if batch else L.nmhd_last_error)(handle)
        raise NmhdError("nmhd error %s: %s" % (ERRORS.get(rc, rc),
                                               msg.decode() if msg else ""))


def desc_from_config(cfg, path=PATH_AUTO):
    d = NmhdDesc()
    for i in range(3):
        d.global_size[i] = cfg["global_size"][i]
        d.margin[i] = cfg["m"][i]
        d.global_extent[i] = cfg["global_extent"][i]
    d.gamma = cfg["gamma"]
    d.cfl_factor = cfg["cfl"]
    d.source = cfg["source"]
    d.omega = cfg["omega"]
    d.qshear = cfg["qshear"]
    d.shear_y = cfg["shear_y"]
    d.recon = cfg.get("recon", 0)
    d.integrator = cfg.get("integrator", 0)
    d.path = path
    d.fixes = cfg.get("fixes", 0)
    return d


class Context:
    """One box on the current HIP device (owns device buffers + a stream)."""

    def __init__(self, cfg, path=PATH_AUTO):
        L = load()
        self.cfg = cfg
        self._desc = desc_from_config(cfg, path)
        h = _VP()
        _check(L.nmhd_create(ctypes.byref(self._desc), ctypes.byref(h)), None)
        self.h = h
        g = NmhdGeometry()
        _check(L.nmhd_get_geometry(self.h, ctypes.byref(g)), self.h)
        self.geometry = g
        self.shape = (g.size[2], g.size[1], g.size[0])
        self.ncell = g.ncell

    def close(self):
        if getattr(self, "h", None):
            load().nmhd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h)

    def set_state(self, W):
        arrs = [np.ascontiguousarray(W[i], dtype=np.float32).reshape(-1) for i in range(8)]
        for a in arrs:
            if a.size != self.ncell:
                raise ValueError("field size %d != box size %d" % (a.size, self.ncell))
        self._c(load().nmhd_set_state(self.h, _ptrs(arrs)))

    def get_state(self):
        out = np.zeros((8,) + self.shape, np.float32)
        self._c(load().nmhd_get_state(self.h, _ptrs([out[i].reshape(-1) for i in range(8)])))
        return out

    def set_state_device(self, ptrs):
        self._c(load().nmhd_set_state_device(self.h, _ptrs(list(ptrs))))

    def get_state_device(self, ptrs):
        self._c(load().nmhd_get_state_device(self.h, _ptrs(list(ptrs))))

    def set_time(self, t, generation=0):
        self._c(load().nmhd_set_time(self.h, t, generation))

    def get_time(self):
        t = ctypes.c_double()
        g = ctypes.c_int()
        self._c(load().nmhd_get_time(self.h, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value

    def fill_ghosts(self):
        self._c(load().nmhd_fill_ghosts(self.h))

    def relative_displacement(self, t):
        nd = ctypes.c_float()
        n = ctypes.c_int()
        mi = ctypes.c_float()
        self._c(load().nmhd_relative_displacement(self.h, t, ctypes.byref(nd),
                                                  ctypes.byref(n), ctypes.byref(mi)))
        return nd.value, n.value, mi.value

    def cfl_dt(self, goal):
        dt = ctypes.c_float()
        self._c(load().nmhd_cfl_dt(self.h, goal, ctypes.byref(dt)))
        return dt.value

    def rhs(self, h):
        out = np.zeros((8,) + self.shape, np.float32)
        self._c(load().nmhd_rhs(self.h, h, _ptrs([out[i].reshape(-1) for i in range(8)])))
        return out

    def step(self, goal=1e30):
        dt = ctypes.c_float()
        r = ctypes.c_int()
        self._c(load().nmhd_step(self.h, goal, ctypes.byref(dt), ctypes.byref(r)))
        return dt.value, bool(r.value)

    def run(self, goal, max_steps):
        k = ctypes.c_int()
        r = ctypes.c_int()
        self._c(load().nmhd_run(self.h, goal, max_steps, ctypes.byref(k), ctypes.byref(r)))
        return k.value, bool(r.value)

    def reset_run(self):
        self._c(load().nmhd_reset_run(self.h))

    def enqueue_steps(self, goal, n):
        self._c(load().nmhd_enqueue_steps(self.h, goal, n))

    def synchronize(self):
        k = ctypes.c_int()
        r = ctypes.c_int()
        self._c(load().nmhd_synchronize(self.h, ctypes.byref(k), ctypes.byref(r)))
        return k.value, bool(r.value)

    def set_path(self, path):
        self._c(load().nmhd_set_path(self.h, path))

    def get_path(self):
        p = ctypes.c_int()
        self._c(load().nmhd_get_path(self.h, ctypes.byref(p)))
        return p.value

    def diagnostics(self):
        """{'b2_mean': [3], 'total': [8], 'divb_max': float} reduced on the GPU."""
        d = NmhdDiagnostics()
        self._c(load().nmhd_get_diagnostics(self.h, ctypes.byref(d)))
        return {"b2_mean": list(d.b2_mean), "total": list(d.total), "divb_max": d.divb_max}

    def set_tiles(self, faces_tile, update_tile):
        self._c(load().nmhd_set_tiles(self.h, faces_tile, update_tile))

    def get_tiles(self):
        f = ctypes.c_int()
        u = ctypes.c_int()
        self._c(load().nmhd_get_tiles(self.h, ctypes.byref(f), ctypes.byref(u)))
        return f.value, u.value

    def timing_enable(self, on=True):
        self._c(load().nmhd_timing_enable(self.h, int(on)))

    def timing_reset(self):
        self._c(load().nmhd_timing_reset(self.h))

    def timing(self):
        """{kernel: (total_ms, launches)} since the last reset."""
        out = {}
        i = 0
        L = load()
        while True:
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            n = ctypes.c_long()
            if L.nmhd_timing_get(self.h, i, ctypes.byref(name), ctypes.byref(ms),
                                 ctypes.byref(n)) != NMHD_OK:
                break
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out


def hlld_batch(inputs, gamma=5.0 / 3.0, strict=False):
    """Device HLLD on (15, n) host face states -> (7, n) fluxes."""
    inputs = np.ascontiguousarray(inputs, dtype=np.float32)
    n = inputs.shape[1]
    ins = [np.ascontiguousarray(inputs[i]) for i in range(15)]
    outs = [np.zeros(n, np.float32) for _ in range(7)]
    _check(load().nmhd_hlld_batch(n, gamma, PATH_STRICT if strict else PATH_AUTO,
                                  _ptrs(ins), _ptrs(outs)))
    return np.stack(outs)


class Batch:
    """Many independent periodic boxes advanced together (nmhd_batch_*):
    per-box time and dt, persistent tile launches, HIP-graph replay."""

    def __init__(self, cfgs):
        L = load()
        self.cfgs = list(cfgs)
        descs = (NmhdDesc * len(self.cfgs))()
        for i, c in enumerate(self.cfgs):
            descs[i] = desc_from_config(c)
        h = _VP()
        _check(L.nmhd_batch_create(descs, len(self.cfgs), ctypes.byref(h)), None, batch=True)
        self.h = h
        self.nbox = len(self.cfgs)
        self.shapes, self.ncells = [], []
        for i in range(self.nbox):
            g = NmhdGeometry()
            self._c(L.nmhd_batch_get_geometry(self.h, i, ctypes.byref(g)))
            self.shapes.append((g.size[2], g.size[1], g.size[0]))
            self.ncells.append(g.ncell)

    def close(self):
        if getattr(self, "h", None):
            load().nmhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.h, batch=True)

    def set_state(self, box, W):
        arrs = [np.ascontiguousarray(W[i], dtype=np.float32).reshape(-1) for i in range(8)]
        for a in arrs:
            if a.size != self.ncells[box]:
                raise ValueError("field size %d != box size %d" % (a.size, self.ncells[box]))
        self._c(load().nmhd_batch_set_state(self.h, box, _ptrs(arrs)))

    def get_state(self, box):
        out = np.zeros((8,) + self.shapes[box], np.float32)
        self._c(load().nmhd_batch_get_state(self.h, box,
                                            _ptrs([out[i].reshape(-1) for i in range(8)])))
        return out

    def set_time(self, box, t, generation=0):
        self._c(load().nmhd_batch_set_time(self.h, box, t, generation))

    def get_time(self, box):
        t = ctypes.c_double()
        g = ctypes.c_int()
        self._c(load().nmhd_batch_get_time(self.h, box, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value

    def status(self, box):
        """(steps taken in the last run, goal reached, last dt)"""
        k = ctypes.c_int()
        r = ctypes.c_int()
        dt = ctypes.c_float()
        self._c(load().nmhd_batch_get_status(self.h, box, ctypes.byref(k), ctypes.byref(r),
                                             ctypes.byref(dt)))
        return k.value, bool(r.value), dt.value

    def set_tiles(self, faces_tile, update_tile, box=-1):
        self._c(load().nmhd_batch_set_tiles(self.h, box, faces_tile, update_tile))

    def set_launch(self, faces_zchunk, update_zchunk, faces_wg_per_cu, update_wg_per_cu):
        self._c(load().nmhd_batch_set_launch(self.h, faces_zchunk, update_zchunk,
                                             faces_wg_per_cu, update_wg_per_cu))

    def get_launch(self):
        v = [ctypes.c_int() for _ in range(4)]
        self._c(load().nmhd_batch_get_launch(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, cache_path=None):
        """Returns the number of box shapes that had to be timed."""
        n = ctypes.c_int()
        p = cache_path.encode() if cache_path else None
        self._c(load().nmhd_batch_autotune(self.h, p, ctypes.byref(n)))
        return n.value

    def run(self, goal, nsteps, use_graph=True):
        """(fewest steps taken by any box, every box reached goal)"""
        k = ctypes.c_int()
        r = ctypes.c_int()
        self._c(load().nmhd_batch_run(self.h, goal, nsteps, int(use_graph), ctypes.byref(k),
                                      ctypes.byref(r)))
        return k.value, bool(r.value)
