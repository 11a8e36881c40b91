"""ctypes bindings of libfwi2d_hip.so (the C ABI declared in include/fwi2d_hip.h).

The library is the product: every call below runs the HIP engine.  Loading fails
loudly when the library has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfwi2d_hip.so")

_lib = None


class Fwi2dGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("nt", C.c_int), ("dt", C.c_double), ("sub", C.c_int),
                ("bc_top", C.c_int), ("bc_bottom", C.c_int), ("bc_left", C.c_int), ("bc_right", C.c_int),
                ("taper_top", C.c_int), ("taper_bottom", C.c_int), ("taper_left", C.c_int),
                ("taper_right", C.c_int), ("taper_strength", C.c_double), ("vti", C.c_int),
                ("gdamp_tmin", C.c_double), ("gdamp_tmax", C.c_double), ("gdamp_power", C.c_double)]


class Fwi2dOpDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("ox", C.c_double), ("oz", C.c_double), ("nt_data", C.c_int), ("dt_data", C.c_double),
                ("nparam", C.c_int), ("ns", C.c_int), ("sxz", C.POINTER(C.c_double)),
                ("nrec", C.POINTER(C.c_int)), ("rxz", C.POINTER(C.c_double)),
                ("courant", C.c_double), ("dt", C.c_double), ("sub", C.c_int),
                ("bc_top", C.c_int), ("bc_bottom", C.c_int), ("bc_left", C.c_int), ("bc_right", C.c_int),
                ("taper_top", C.c_int), ("taper_bottom", C.c_int), ("taper_left", C.c_int),
                ("taper_right", C.c_int), ("taper_strength", C.c_double), ("seismotype", C.c_int),
                ("resampling", C.c_int), ("sinc_half_length", C.c_int),
                ("vpmin", C.c_double), ("vpmax", C.c_double), ("vsmin", C.c_double), ("vsmax", C.c_double),
                ("rhomin", C.c_double), ("rhomax", C.c_double), ("deltamin", C.c_double),
                ("deltamax", C.c_double), ("epsilonmin", C.c_double), ("epsilonmax", C.c_double),
                ("soft_clip", C.c_int), ("gdamp_tmin", C.c_double), ("gdamp_tmax", C.c_double),
                ("gdamp_power", C.c_double), ("device", C.c_int), ("mt", C.c_int), ("gl", C.c_double),
                ("rdip", C.POINTER(C.c_double)), ("free_surface_stiffness", C.c_double)]


EXPORTS = [
    "fwi2d_last_error", "fwi2d_engine_create", "fwi2d_engine_set_model", "fwi2d_engine_set_shots",
    "fwi2d_engine_forward", "fwi2d_engine_adjoint", "fwi2d_engine_dot_f64", "fwi2d_engine_set_option",
    "fwi2d_engine_get_stat", "fwi2d_engine_tune_report", "fwi2d_engine_destroy", "fwi2d_op_create", "fwi2d_op_apply_forward",
    "fwi2d_op_apply_jacobianT", "fwi2d_op_engine", "fwi2d_op_time_axis", "fwi2d_op_destroy",
    "fwi2d_nlls_create", "fwi2d_nlls_evaluate", "fwi2d_nlls_residual", "fwi2d_nlls_dnorm", "fwi2d_nlls_destroy",
    "fwi2d_nlls_objective", "fwi2d_lbfgs", "fwi2d_nlls_create_filtered", "fwi2d_filter_prepare",
    "fwi2d_filter_apply", "fwi2d_dot_f64_noalloc", "fwi2d_read_header", "fwi2d_read_data", "fwi2d_write",
    "fwi2d_read_coord", "fwi2d_regular_geometry", "fwi2d_engine_set_acquisition", "fwi2d_acoustic_create",
    "fwi2d_acoustic_set_model", "fwi2d_acoustic_set_shots", "fwi2d_acoustic_forward", "fwi2d_acoustic_adjoint",
    "fwi2d_acoustic_set_option", "fwi2d_acoustic_get_stat", "fwi2d_acoustic_destroy",
]

FORMAT_SEP, FORMAT_BINARY = 0, 1

FILTER_PHASES = {"default": 0, "zero": 1, "minimum": 2}

OBJECTIVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_float))


class LbfgsOptions(C.Structure):
    _fields_ = [("niter", C.c_int), ("max_trial", C.c_int), ("threshold", C.c_double), ("memory", C.c_int),
                ("c1", C.c_double), ("c2", C.c_double), ("a0", C.c_double), ("a1", C.c_double),
                ("step_max", C.c_double), ("version", C.c_int), ("H0", C.c_void_p)]


class LbfgsReport(C.Structure):
    _fields_ = [("feval", C.c_int), ("geval", C.c_int), ("ntrials", C.c_int), ("trials_cap", C.c_int),
                ("trials", C.c_void_p)]


class Hypercube(C.Structure):
    _fields_ = [("ndim", C.c_int), ("n", C.c_longlong * 6), ("o", C.c_double * 6), ("d", C.c_double * 6),
                ("label", (C.c_char * 64) * 6), ("esize", C.c_int), ("xdr", C.c_int), ("in_", C.c_char * 4096)]


class Fwi2dError(RuntimeError):
    pass


def load():
    """Load the HIP library (built by fwi2d_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Fwi2dError("%s is missing: build it with `python -m fwi2d_amd.build`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.fwi2d_last_error.restype = C.c_char_p
    lib.fwi2d_engine_create.argtypes = [C.POINTER(Fwi2dGrid), C.c_int, C.POINTER(P)]
    lib.fwi2d_engine_set_model.argtypes = [P, P]
    lib.fwi2d_engine_set_shots.argtypes = [P, C.c_int, P, P, P, C.c_double, C.c_double]
    lib.fwi2d_engine_set_acquisition.argtypes = [P, C.c_int, P, P, P, C.c_double, C.c_double, C.c_int, C.c_double, P]
    lib.fwi2d_engine_forward.argtypes = [P, P, P, C.c_int, C.c_int]
    lib.fwi2d_engine_adjoint.argtypes = [P, P, P, C.c_int]
    lib.fwi2d_engine_dot_f64.argtypes = [P, P, P, P, C.c_longlong, C.POINTER(C.c_double)]
    lib.fwi2d_dot_f64_noalloc.argtypes = [C.c_int, P, P, P, C.c_longlong, C.POINTER(C.c_double)]
    lib.fwi2d_engine_set_option.argtypes = [P, C.c_char_p, C.c_longlong]
    lib.fwi2d_engine_get_stat.argtypes = [P, C.c_char_p]
    lib.fwi2d_engine_get_stat.restype = C.c_longlong
    lib.fwi2d_engine_tune_report.argtypes = [P]
    lib.fwi2d_engine_tune_report.restype = C.c_char_p
    lib.fwi2d_engine_destroy.argtypes = [P]
    lib.fwi2d_acoustic_create.argtypes = [C.POINTER(Fwi2dGrid), C.c_double, C.c_int, C.POINTER(P)]
    lib.fwi2d_acoustic_set_model.argtypes = [P, P]
    lib.fwi2d_acoustic_set_shots.argtypes = [P, C.c_int, P, P, P, C.c_double, C.c_double]
    lib.fwi2d_acoustic_forward.argtypes = [P, P, P, C.c_int, C.c_int]
    lib.fwi2d_acoustic_adjoint.argtypes = [P, P, P, C.c_int]
    lib.fwi2d_acoustic_set_option.argtypes = [P, C.c_char_p, C.c_longlong]
    lib.fwi2d_acoustic_get_stat.argtypes = [P, C.c_char_p]
    lib.fwi2d_acoustic_get_stat.restype = C.c_longlong
    lib.fwi2d_acoustic_destroy.argtypes = [P]
    lib.fwi2d_op_create.argtypes = [C.POINTER(Fwi2dOpDesc), P, C.POINTER(P)]
    lib.fwi2d_op_apply_forward.argtypes = [P, C.c_int, P, P]
    lib.fwi2d_op_apply_jacobianT.argtypes = [P, C.c_int, P, P, P]
    lib.fwi2d_op_engine.argtypes = [P]
    lib.fwi2d_op_engine.restype = P
    lib.fwi2d_op_time_axis.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.fwi2d_op_destroy.argtypes = [P]
    lib.fwi2d_nlls_create.argtypes = [P, P, P, P, C.c_int, C.POINTER(P)]
    lib.fwi2d_nlls_create_filtered.argtypes = [P, P, P, P, C.c_int, P, C.c_int, C.c_int, C.POINTER(P)]
    lib.fwi2d_filter_prepare.argtypes = [P, C.c_int, C.c_int, P, C.POINTER(C.c_int)]
    lib.fwi2d_filter_apply.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_int, P, P]
    lib.fwi2d_nlls_evaluate.argtypes = [P, P, C.POINTER(C.c_double), P]
    lib.fwi2d_nlls_residual.argtypes = [P]
    lib.fwi2d_nlls_residual.restype = C.POINTER(C.c_float)
    lib.fwi2d_nlls_dnorm.argtypes = [P]
    lib.fwi2d_nlls_dnorm.restype = C.c_double
    lib.fwi2d_nlls_destroy.argtypes = [P]
    lib.fwi2d_read_header.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(Hypercube)]
    lib.fwi2d_read_data.argtypes = [C.POINTER(Hypercube), P]
    lib.fwi2d_write.argtypes = [C.c_char_p, C.c_int, C.POINTER(Hypercube), P, C.c_char_p]
    lib.fwi2d_read_coord.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong), P, P, P]
    lib.fwi2d_regular_geometry.argtypes = [C.c_int] + [C.c_float] * 4 + [C.c_int] + [C.c_float] * 6 + [P, P]
    lib.fwi2d_lbfgs.argtypes = [C.c_void_p, P, C.c_longlong, P, C.POINTER(LbfgsOptions), P, C.POINTER(C.c_int),
                                C.POINTER(LbfgsReport)]
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise Fwi2dError(load().fwi2d_last_error().decode())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def grid_from(prob, nt, dt, sub):
    g = Fwi2dGrid()
    for k in ("nx", "nz", "dx", "dz", "bc_top", "bc_bottom", "bc_left", "bc_right", "taper_top", "taper_bottom",
              "taper_left", "taper_right", "taper_strength", "vti", "gdamp_tmin", "gdamp_tmax", "gdamp_power"):
        if k in prob:
            setattr(g, k, prob[k])
    g.nt, g.dt, g.sub = int(nt), float(dt), int(sub)
    return g


class AcousticEngine:
    """Batched device propagation of the acoustic operator (nl_we_op_a), engine level of
    the C ABI: model (K, 1/rho) [2][nx][nz], one trace component."""

    def __init__(self, grid, free_surface_stiffness=1.05, device=0):
        lib = load()
        self._h = C.c_void_p()
        self.grid = grid
        _check(lib.fwi2d_acoustic_create(C.byref(grid), float(free_surface_stiffness), int(device),
                                         C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            load().fwi2d_acoustic_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def set_model(self, model):
        m = _f32(model)
        assert m.shape == (2, self.grid.nx, self.grid.nz), m.shape
        _check(load().fwi2d_acoustic_set_model(self._h, _ptr(m)))

    def set_shots(self, sxz, rxz, ox=0.0, oz=0.0):
        sxz = np.ascontiguousarray(sxz, dtype=np.float64).reshape(-1, 2)
        nrec = np.ascontiguousarray([len(r) for r in rxz], dtype=np.int32)
        allr = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64).reshape(-1, 2) for r in rxz], 0))
        self.ns, self.nr = len(sxz), int(nrec.sum())
        self._geom = (sxz, nrec, allr)
        _check(load().fwi2d_acoustic_set_shots(self._h, self.ns, _ptr(sxz), _ptr(nrec), _ptr(allr), ox, oz))

    def forward(self, src, keep_snapshots=False):
        """src [ns][nt] (host numpy) -> receivers [nr][nt]."""
        s = _f32(src)
        assert s.shape == (self.ns, self.grid.nt), s.shape
        out = np.zeros((self.nr, self.grid.nt), np.float32)
        _check(load().fwi2d_acoustic_forward(self._h, _ptr(s), _ptr(out), int(keep_snapshots), 0))
        return out

    def adjoint(self, adj):
        """adjoint sources [nr][nt] -> raw (K, 1/rho) gradient [2][nx][nz]."""
        a = _f32(adj)
        assert a.shape == (self.nr, self.grid.nt), a.shape
        g = np.zeros((2, self.grid.nx, self.grid.nz), np.float32)
        _check(load().fwi2d_acoustic_adjoint(self._h, _ptr(a), _ptr(g), 0))
        return g

    def set_option(self, key, value):
        _check(load().fwi2d_acoustic_set_option(self._h, key.encode(), int(value)))

    def stat(self, key):
        return load().fwi2d_acoustic_get_stat(self._h, key.encode())

    def forward_device(self, src_ptr, rcv_ptr, keep_snapshots=False):
        """Device pointers: src [ns][nt] -> rcv [nr][nt]."""
        _check(load().fwi2d_acoustic_forward(self._h, C.c_void_p(src_ptr), C.c_void_p(rcv_ptr),
                                             int(keep_snapshots), 1))

    def adjoint_device(self, adj_ptr, grad_ptr):
        """Device pointers: adjoint sources [nr][nt] -> raw gradient [2][nx][nz]."""
        _check(load().fwi2d_acoustic_adjoint(self._h, C.c_void_p(adj_ptr), C.c_void_p(grad_ptr), 1))


class Engine:
    """Batched device propagation (engine level of the C ABI)."""

    def __init__(self, grid, device=0):
        lib = load()
        self._h = C.c_void_p()
        self.grid = grid
        _check(lib.fwi2d_engine_create(C.byref(grid), int(device), C.byref(self._h)))
        self.np = 5 if grid.vti else 3

    def close(self):
        if getattr(self, "_h", None):
            load().fwi2d_engine_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def set_model(self, model):
        m = _f32(model)
        assert m.shape == (self.np, self.grid.nx, self.grid.nz), m.shape
        _check(load().fwi2d_engine_set_model(self._h, _ptr(m)))

    def set_shots(self, sxz, rxz, ox=0.0, oz=0.0, moment_tensor=False, gauge_length=0.0, rdip=None):
        """Point forces (or moment tensors) at sxz, point receivers (or DAS channels of
        gauge_length along the dips rdip, one array per shot) at rxz (one array per shot)."""
        sxz = np.ascontiguousarray(sxz, dtype=np.float64).reshape(-1, 2)
        nrec = np.ascontiguousarray([len(r) for r in rxz], dtype=np.int32)
        allr = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64).reshape(-1, 2) for r in rxz], 0))
        self.ns, self.nr = len(sxz), int(nrec.sum())
        self.src_nc = 3 if moment_tensor else 2
        dips = None
        if gauge_length > 0:
            dips = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64) for r in rdip]))
        self._geom = (sxz, nrec, allr, dips)
        _check(load().fwi2d_engine_set_acquisition(self._h, self.ns, _ptr(sxz), _ptr(nrec), _ptr(allr), ox, oz,
                                                   int(moment_tensor), float(gauge_length), _ptr(dips)))

    def forward(self, src, keep_checkpoints=False):
        """src [2][ns][nt] (host numpy) -> receivers [2][nr][nt]."""
        s = _f32(src)
        assert s.shape == (getattr(self, "src_nc", 2), self.ns, self.grid.nt), s.shape
        rcv = np.zeros((2, self.nr, self.grid.nt), np.float32)
        _check(load().fwi2d_engine_forward(self._h, _ptr(s), _ptr(rcv), int(keep_checkpoints), 0))
        return rcv

    def forward_device(self, src_ptr, rcv_ptr, keep_checkpoints=False):
        _check(load().fwi2d_engine_forward(self._h, C.c_void_p(src_ptr), C.c_void_p(rcv_ptr),
                                           int(keep_checkpoints), 1))

    def adjoint(self, adj):
        a = _f32(adj)
        assert a.shape == (2, self.nr, self.grid.nt), a.shape
        g = np.zeros((self.np, self.grid.nx, self.grid.nz), np.float32)
        _check(load().fwi2d_engine_adjoint(self._h, _ptr(a), _ptr(g), 0))
        return g

    def adjoint_device(self, adj_ptr, grad_ptr):
        _check(load().fwi2d_engine_adjoint(self._h, C.c_void_p(adj_ptr), C.c_void_p(grad_ptr), 1))

    def dot_f64(self, a_ptr, b_ptr, w_ptr, n):
        out = C.c_double()
        _check(load().fwi2d_engine_dot_f64(self._h, C.c_void_p(a_ptr), C.c_void_p(b_ptr) if b_ptr else None,
                                           C.c_void_p(w_ptr) if w_ptr else None, int(n), C.byref(out)))
        return out.value

    def set_option(self, key, value):
        _check(load().fwi2d_engine_set_option(self._h, key.encode(), int(value)))

    def stat(self, key):
        return load().fwi2d_engine_get_stat(self._h, key.encode())

    def tune_report(self):
        return load().fwi2d_engine_tune_report(self._h).decode()


def op_desc_from(prob, device=0):
    """Operator description from a problem dict (see fwi2d_amd.workloads)."""
    d = Fwi2dOpDesc()
    mp = prob["mp"]
    for k in ("nx", "nz", "dx", "dz", "ox", "oz", "bc_top", "bc_bottom", "bc_left", "bc_right", "taper_top",
              "taper_bottom", "taper_left", "taper_right", "taper_strength", "seismotype", "resampling",
              "gdamp_tmin", "gdamp_tmax", "gdamp_power"):
        if k in prob:
            setattr(d, k, prob[k])
    d.sinc_half_length = int(prob.get("sinc_hl", 11))
    d.nparam = 2 if prob.get("acoustic") else (5 if prob.get("vti") else 3)
    d.free_surface_stiffness = float(prob.get("free_surface_stiffness", 1.05))  # param.hpp:35
    d.nt_data, d.dt_data = int(mp["nt_data"]), float(mp["dt_data"])
    d.courant, d.dt, d.sub = float(mp["courant"]), float(mp["dt"]), int(mp["sub"])
    d.vpmin, d.vpmax, d.vsmin, d.vsmax = mp["vpmin"], mp["vpmax"], mp["vsmin"], mp["vsmax"]
    d.rhomin, d.rhomax = mp["rhomin"], mp["rhomax"]
    d.deltamin, d.deltamax, d.epsilonmin, d.epsilonmax = mp["deltamin"], mp["deltamax"], mp["epsmin"], mp["epsmax"]
    d.soft_clip = int(mp.get("soft_clip", 0))
    d.device = int(device)
    return d


class WeOp:
    """Drop-in for the reference's nl_we_op_e / nl_we_op_vti / nl_we_op_a: apply_forward
    and apply_jacobianT on host numpy arrays, computed by the HIP engine.  prob["acoustic"]
    selects nl_we_op_a: model (vp, rho), sources [1][ns][nt_data], data [1][nr][nt_data]."""

    def __init__(self, prob, allsrc, device=0):
        lib = load()
        self.prob = prob
        self.acoustic = bool(prob.get("acoustic", 0))
        self.np = 2 if self.acoustic else (5 if prob.get("vti") else 3)
        self.ncd = 1 if self.acoustic else 2
        d = op_desc_from(prob, device)
        sxz = np.ascontiguousarray(prob["sxz"], dtype=np.float64)
        nrec = np.ascontiguousarray([len(r) for r in prob["rxz"]], dtype=np.int32)
        rxz = np.ascontiguousarray(np.concatenate(prob["rxz"], 0), dtype=np.float64)
        d.ns = len(sxz)
        d.sxz = sxz.ctypes.data_as(C.POINTER(C.c_double))
        d.nrec = nrec.ctypes.data_as(C.POINTER(C.c_int))
        d.rxz = rxz.ctypes.data_as(C.POINTER(C.c_double))
        self.nr = int(nrec.sum())
        # the acoustic operator has point sources and receivers only (we_op.cpp:13-20)
        d.mt = 0 if self.acoustic else int(prob.get("mt", 0))
        d.gl = 0.0 if self.acoustic else float(prob.get("gl", 0.0))
        if d.gl > 0:
            self._rdip = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float64) for r in prob["rdip"]]))
            assert self._rdip.size == self.nr
            d.rdip = self._rdip.ctypes.data_as(C.POINTER(C.c_double))
        src = _f32(allsrc)
        nc = 1 if self.acoustic else (3 if d.mt else 2)
        assert src.shape == (nc, d.ns, d.nt_data), src.shape
        self._h = C.c_void_p()
        _check(lib.fwi2d_op_create(C.byref(d), _ptr(src), C.byref(self._h)))
        self.nt_data = d.nt_data

    def close(self):
        if getattr(self, "_h", None):
            load().fwi2d_op_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def apply_forward(self, model, add_to=None):
        m = _f32(model)
        out = np.zeros((self.ncd, self.nr, self.nt_data), np.float32) if add_to is None else _f32(add_to).copy()
        _check(load().fwi2d_op_apply_forward(self._h, int(add_to is not None), _ptr(m), _ptr(out)))
        return out

    def apply_jacobianT(self, model0, data, add_to=None):
        m0 = _f32(model0)
        d = _f32(data)
        g = np.zeros((self.np,) + m0.shape[1:], np.float32) if add_to is None else _f32(add_to).copy()
        _check(load().fwi2d_op_apply_jacobianT(self._h, int(add_to is not None), _ptr(g), _ptr(m0), _ptr(d)))
        return g

    def time_axis(self):
        nt, dt, sub = C.c_int(), C.c_double(), C.c_int()
        _check(load().fwi2d_op_time_axis(self._h, C.byref(nt), C.byref(dt), C.byref(sub)))
        return nt.value, dt.value, sub.value

    def engine_handle(self):
        return load().fwi2d_op_engine(self._h)


NLLS_NORMALIZE_OBJ, NLLS_INTEGRATE, NLLS_TRACE_NORMALIZE = 1, 2, 4


class Nlls:
    """nlls_fwi (normalize=False) / nlls_fwi_eco (normalize=True: the functional divided
    by dt |P d|^2) on a WeOp, with residual weights, time integration, a 1-D filter
    (filter_phase "default", "zero" or "minimum", as par.filter_phase) and trace
    normalisation as options: evaluate(m) returns the functional (FP64) and the gradient."""

    def __init__(self, op, d, w=None, gmask=None, normalize=False, integrate=False, trace_normalize=False,
                 filter=None, filter_phase="default"):
        self.op = op  # the objective uses the operator handle
        self._d = _f32(d)
        assert self._d.shape == (op.ncd, op.nr, op.nt_data), self._d.shape
        w = None if w is None else _f32(w)
        gmask = None if gmask is None else _f32(gmask)
        filt = None if filter is None else _f32(filter).ravel()
        self._h = C.c_void_p()
        flags = (NLLS_NORMALIZE_OBJ * bool(normalize) | NLLS_INTEGRATE * bool(integrate)
                 | NLLS_TRACE_NORMALIZE * bool(trace_normalize))
        _check(load().fwi2d_nlls_create_filtered(op._h, _ptr(self._d), _ptr(w), _ptr(gmask), flags, _ptr(filt),
                                                 0 if filt is None else filt.size, FILTER_PHASES[filter_phase],
                                                 C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            load().fwi2d_nlls_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def evaluate(self, model, gradient=True):
        m = _f32(model)
        f = C.c_double()
        g = np.zeros((self.op.np,) + m.shape[1:], np.float32) if gradient else None
        _check(load().fwi2d_nlls_evaluate(self._h, _ptr(m), C.byref(f), _ptr(g)))
        return f.value, g

    def residual(self):
        p = load().fwi2d_nlls_residual(self._h)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=self._d.shape).copy()

    def dnorm(self):
        return load().fwi2d_nlls_dnorm(self._h)


def filter_prepare(f, phase="default"):
    """The objective's filter preparation (zero_phase / minimum_phase): (filter, centred)."""
    f = _f32(f).ravel()
    out = np.zeros(f.size + 1)
    n = C.c_int()
    _check(load().fwi2d_filter_prepare(_ptr(f), f.size, FILTER_PHASES[phase], _ptr(out), C.byref(n)))
    return out[:n.value].copy(), phase != "minimum"


def filter_apply(f, traces, centered=True, adjoint=False):
    """conv1dnd forward (or adjoint) of every trace (last axis) with the filter f, in double precision."""
    f = np.ascontiguousarray(f, np.float64).ravel()
    x = np.ascontiguousarray(traces, np.float64)
    out = np.zeros_like(x)
    nt = x.shape[-1]
    _check(load().fwi2d_filter_apply(_ptr(f), f.size, int(centered), int(adjoint), x.size // nt, nt, _ptr(x),
                                     _ptr(out)))
    return out


def lbfgs(objective, m0, niter=10, max_trial=10, threshold=0.0, memory=5, c1=1e-4, c2=0.9, a0=1.0, a1=0.0,
          step_max=1e6, version=0, H0=None, report=False):
    """The reference's l-BFGS with the regular Wolfe line search (fwi2d_lbfgs).  objective:
    an Nlls (run through fwi2d_nlls_objective) or a Python callable m -> (f, g).
    Returns (model, functional history, iterations), and with report=True also a dict
    with the evaluation counts ("feval", "geval") and the line-search trials ("trials":
    rows of iteration, trial, step length, function value, gradient norm or NaN)."""
    lib = load()
    m = _f32(m0).copy()
    n = m.size
    h0 = None if H0 is None else _f32(H0).ravel()
    opt = LbfgsOptions(int(niter), int(max_trial), float(threshold), int(memory), float(c1), float(c2), float(a0),
                       float(a1), float(step_max), int(version), None if h0 is None else h0.ctypes.data)
    hist = np.zeros(niter + 1)
    iters = C.c_int()
    cap = max(1, niter * max_trial)
    rows = np.zeros((cap, 5))
    rep = LbfgsReport(0, 0, 0, cap, rows.ctypes.data)
    errors = []
    if isinstance(objective, Nlls):
        fn = C.cast(lib.fwi2d_nlls_objective, C.c_void_p)
        ctx = objective._h
    else:
        def call(_ctx, pm, pf, pg):
            try:
                f, g = objective(np.ctypeslib.as_array(pm, shape=(n,)).copy().reshape(m.shape))
                pf[0] = float(f)
                np.ctypeslib.as_array(pg, shape=(n,))[:] = np.asarray(g, np.float32).ravel()
                return 0
            except Exception as ex:  # the solver stops with an error; re-raised below
                errors.append(ex)
                return 1
        cb = OBJECTIVE_FN(call)
        fn = C.cast(cb, C.c_void_p)
        ctx = None
    rc = lib.fwi2d_lbfgs(fn, ctx, n, _ptr(m), C.byref(opt), _ptr(hist), C.byref(iters), C.byref(rep))
    if rc != 0 and errors:
        raise Fwi2dError(lib.fwi2d_last_error().decode()) from errors[0]
    _check(rc)
    k = iters.value
    if report:
        info = dict(feval=rep.feval, geval=rep.geval, trials=rows[:min(rep.ntrials, cap)].copy())
        return m, hist[:k + 1], k, info
    return m, hist[:k + 1], k


# ---------------------------------------------------------------------------
# On-disk formats (host only)
# ---------------------------------------------------------------------------
def read_header(path, format=FORMAT_SEP, ndim0=1):
    """The header of a SEPlib (format 0) or binary-descriptor (format 1) file, as a dict
    with n, o, d, label (one entry per kept axis, axis 1 first), esize, xdr and in."""
    h = Hypercube()
    _check(load().fwi2d_read_header(os.fsencode(path), int(format), int(ndim0), C.byref(h)))
    nd = h.ndim
    return dict(n=[h.n[i] for i in range(nd)], o=[h.o[i] for i in range(nd)], d=[h.d[i] for i in range(nd)],
                label=[h.label[i].value.decode() for i in range(nd)], esize=h.esize, xdr=h.xdr,
                in_=h.in_.decode(), _h=h)


def read(path, format=FORMAT_SEP, ndim0=1):
    """read<data_t>(file, format) of the reference: (array, header) with the array in
    C order [nN]...[n1] (axis 1 fastest) and single precision."""
    hd = read_header(path, format, ndim0)
    h = hd.pop("_h")
    n = [h.n[i] for i in range(6)]
    out = np.zeros(int(np.prod(n)), np.float32)
    _check(load().fwi2d_read_data(C.byref(h), _ptr(out)))
    return out.reshape(hd["n"][::-1]), hd


def write(path, data, o=None, d=None, label=None, format=FORMAT_SEP, datapath=None):
    """write<data_t>(vec, file, format) of the reference: data in C order [nN]...[n1]
    (axis 1 fastest), o, d, label per axis from axis 1 (defaults 0, 1, none)."""
    a = _f32(data)
    n = list(a.shape[::-1]) or [1]
    if len(n) > 6:
        raise ValueError("at most 6 axes")
    h = Hypercube()
    h.ndim = len(n)
    for i, ni in enumerate(n):
        h.n[i] = ni
        h.o[i] = 0.0 if o is None else float(o[i])
        h.d[i] = 1.0 if d is None else float(d[i])
        if label is not None and label[i]:
            h.label[i].value = str(label[i]).encode()[:63]
    _check(load().fwi2d_write(os.fsencode(path), int(format), C.byref(h), _ptr(a),
                              None if datapath is None else os.fsencode(datapath)))


def read_coord(path):
    """readCoord: (sxz [ns][2], rxz list of [nrec][2] arrays, dips list of [nrec] arrays)."""
    lib = load()
    ns, nr = C.c_int(0), C.c_longlong(0)
    _check(lib.fwi2d_read_coord(os.fsencode(path), C.byref(ns), C.byref(nr), None, None, None))
    sxz = np.zeros((ns.value, 2))
    nrec = np.zeros(ns.value, np.int32)
    rxz = np.zeros((nr.value, 3))
    _check(lib.fwi2d_read_coord(os.fsencode(path), C.byref(ns), C.byref(nr), _ptr(sxz), _ptr(nrec), _ptr(rxz)))
    off = np.concatenate([[0], np.cumsum(nrec)])
    return sxz, [rxz[off[i]:off[i + 1], :2].copy() for i in range(ns.value)], \
        [rxz[off[i]:off[i + 1], 2].copy() for i in range(ns.value)]


def regular_geometry(ns, sx0, sz0, sxinc, szinc, nr, rx0, rz0, rxinc, rzinc, rdip0=0.0, rdipinc=0.0):
    """The regular acquisition of the shot parameters: (sxz [ns][2], rxz [ns][nr][3])."""
    sxz = np.zeros((ns, 2))
    rxz = np.zeros((ns, nr, 3))
    _check(load().fwi2d_regular_geometry(ns, sx0, sz0, sxinc, szinc, nr, rx0, rz0, rxinc, rzinc, rdip0, rdipinc,
                                         _ptr(sxz), _ptr(rxz)))
    return sxz, rxz
