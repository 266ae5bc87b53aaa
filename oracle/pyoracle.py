"""ctypes binding of the oracle (test infrastructure only).

Loaded by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py as the checker.  The product path (nmhd_amd) never imports this.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

FIELDS = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")


class OrGrid(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("m", ctypes.c_int * 3),
        ("d", ctypes.c_float * 3),
        ("ext", ctypes.c_float * 3),
        ("local_ext", ctypes.c_float * 3),
        ("gamma", ctypes.c_float),
        ("cfl_factor", ctypes.c_float),
        ("source", ctypes.c_int),
        ("omega", ctypes.c_float),
        ("qshear", ctypes.c_float),
        ("shear_y", ctypes.c_int),
        ("recon", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("fixes", ctypes.c_int),
    ]


_FP = ctypes.POINTER(ctypes.c_float)
_FPP = ctypes.POINTER(_FP)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `python oracle/build.py`")
        L = ctypes.CDLL(path)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.POINTER(OrGrid)]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_set_threads.argtypes = [ctypes.c_int]
        for name in ("or_set_state", "or_get_state", "or_get_predictor"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _FPP]
        L.or_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.or_get_time.restype = ctypes.c_double
        L.or_get_time.argtypes = [ctypes.c_void_p]
        L.or_get_generation.argtypes = [ctypes.c_void_p]
        L.or_relative_displacement.argtypes = [
            ctypes.POINTER(OrGrid), ctypes.c_double, _FP,
            ctypes.POINTER(ctypes.c_int), _FP]
        L.or_fill_ghosts.argtypes = [ctypes.c_void_p]
        L.or_cfl_dt.restype = ctypes.c_float
        L.or_cfl_dt.argtypes = [ctypes.c_void_p, ctypes.c_float]
        L.or_cfl_field.argtypes = [ctypes.c_void_p, ctypes.c_float, _FP]
        L.or_rhs.argtypes = [ctypes.c_void_p, ctypes.c_float, _FPP]
        L.or_face_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, _FPP]
        L.or_step.restype = ctypes.c_float
        L.or_step.argtypes = [ctypes.c_void_p, ctypes.c_float,
                              ctypes.POINTER(ctypes.c_int)]
        L.or_run.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int,
                             ctypes.POINTER(ctypes.c_int)]
        L.or_hlld_batch.argtypes = [ctypes.c_size_t, ctypes.c_float, _FPP, _FPP]
        L.or_interpolate_batch.argtypes = [ctypes.c_size_t, ctypes.c_int] + [_FP] * 6
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own hlld.inl built for the host, or None if absent."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_hlld.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_hlld_batch.argtypes = [ctypes.c_size_t, ctypes.c_float, _FPP, _FPP]
        _REF = L
    return _REF


def _ptrs(arrays):
    arr = (_FP * len(arrays))()
    for i, a in enumerate(arrays):
        assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
        arr[i] = a.ctypes.data_as(_FP)
    return arr


def hlld_batch(inputs, gamma=5.0 / 3.0, use_ref=False):
    """inputs: (15, n) float32 -> (7, n) float32 fluxes."""
    inputs = np.ascontiguousarray(inputs, dtype=np.float32)
    n = inputs.shape[1]
    out = np.zeros((7, n), dtype=np.float32)
    ins = [np.ascontiguousarray(inputs[i]) for i in range(15)]
    outs = [np.ascontiguousarray(out[i]) for i in range(7)]
    L = ref_lib() if use_ref else lib()
    if L is None:
        raise RuntimeError("reference HLLD build (oracle/_ref) not present")
    fn = L.ref_hlld_batch if use_ref else L.or_hlld_batch
    fn(n, ctypes.c_float(gamma), _ptrs(ins), _ptrs(outs))
    return np.stack(outs)


def interpolate_batch(x, recon=0):
    """x: (4, n) float32 -> (left, right)."""
    x = [np.ascontiguousarray(v, dtype=np.float32) for v in x]
    n = x[0].size
    left = np.zeros(n, np.float32)
    right = np.zeros(n, np.float32)
    lib().or_interpolate_batch(n, recon, *[v.ctypes.data_as(_FP) for v in x],
                               left.ctypes.data_as(_FP), right.ctypes.data_as(_FP))
    return left, right


def make_grid(cfg):
    """OrGrid from a nmhd_amd-style config dict (see nmhd_amd.problems)."""
    g = OrGrid()
    for i in range(3):
        g.n[i] = cfg["n"][i]
        g.m[i] = cfg["m"][i]
        g.d[i] = cfg["d"][i]
        g.ext[i] = cfg["ext"][i]
        g.local_ext[i] = cfg["local_ext"][i]
    g.gamma = cfg["gamma"]
    g.cfl_factor = cfg["cfl"]
    g.source = cfg["source"]
    g.omega = cfg["omega"]
    g.qshear = cfg["qshear"]
    g.shear_y = cfg["shear_y"]
    g.recon = cfg.get("recon", 0)
    g.integrator = cfg.get("integrator", 0)
    g.fixes = cfg.get("fixes", 0)
    return g


def relative_displacement(cfg, t):
    g = make_grid(cfg)
    nd = ctypes.c_float()
    n = ctypes.c_int()
    minor = ctypes.c_float()
    lib().or_relative_displacement(ctypes.byref(g), t, ctypes.byref(nd),
                                   ctypes.byref(n), ctypes.byref(minor))
    return nd.value, n.value, minor.value


class Oracle:
    """The reference step on the CPU, for one box (array incl. ghosts)."""

    def __init__(self, cfg, state=None, threads=0):
        self.cfg = cfg
        self.grid = make_grid(cfg)
        L = lib()
        if threads:
            L.or_set_threads(threads)
        self.h = L.or_create(ctypes.byref(self.grid))
        if not self.h:
            raise ValueError("or_create rejected the grid")
        self.shape = (cfg["n"][2], cfg["n"][1], cfg["n"][0])
        if state is not None:
            self.set_state(state)

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    __del__ = close

    def set_state(self, state):
        arrs = [np.ascontiguousarray(state[i], dtype=np.float32).reshape(-1)
                for i in range(8)]
        lib().or_set_state(self.h, _ptrs(arrs))

    def get_state(self, predictor=False):
        out = [np.zeros(self.shape, np.float32) for _ in range(8)]
        fn = lib().or_get_predictor if predictor else lib().or_get_state
        fn(self.h, _ptrs([o.reshape(-1) for o in out]))
        return np.stack(out)

    def set_time(self, t, generation=0):
        lib().or_set_time(self.h, t, generation)

    @property
    def time(self):
        return lib().or_get_time(self.h)

    @property
    def generation(self):
        return lib().or_get_generation(self.h)

    def fill_ghosts(self):
        lib().or_fill_ghosts(self.h)

    def cfl_dt(self, goal):
        return lib().or_cfl_dt(self.h, goal)

    def cfl_field(self, goal):
        out = np.zeros(self.shape, np.float32)
        lib().or_cfl_field(self.h, goal, out.ctypes.data_as(_FP))
        return out

    def rhs(self, h):
        out = [np.zeros(self.shape, np.float32) for _ in range(8)]
        lib().or_rhs(self.h, h, _ptrs([o.reshape(-1) for o in out]))
        return np.stack(out)

    def face_flux(self, d, h):
        """(7,) + shape fluxes through the lower d-faces of the current
        state: mass, momentum x, y, z, energy, B_{d+1}, B_{d+2}."""
        out = [np.zeros(self.shape, np.float32) for _ in range(7)]
        lib().or_face_flux(self.h, d, h, _ptrs([o.reshape(-1) for o in out]))
        return np.stack(out)

    def step(self, goal=1e30):
        reached = ctypes.c_int()
        dt = lib().or_step(self.h, goal, ctypes.byref(reached))
        return dt, bool(reached.value)

    def run(self, goal, max_steps):
        reached = ctypes.c_int()
        k = lib().or_run(self.h, goal, max_steps, ctypes.byref(reached))
        return k, bool(reached.value)
