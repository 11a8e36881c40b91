"""ctypes wrapper of the fwi2d oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
bench.py, as the checker.  The product package (fwi2d_amd) never imports it.

`Oracle(64)` wraps liboracle64.so (REAL=double), `Oracle(32)` liboracle32.so
(REAL=float).  `RefSBP()` wraps _ref/libref_sbp.so, the reference's own SBP
templates compiled from source (present only where the reference was).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("ox", C.c_double), ("oz", C.c_double), ("nt", C.c_int), ("dt", C.c_double),
                ("sub", C.c_int), ("bc_top", C.c_int), ("bc_bottom", C.c_int), ("bc_left", C.c_int),
                ("bc_right", C.c_int), ("taper_top", C.c_int), ("taper_bottom", C.c_int),
                ("taper_left", C.c_int), ("taper_right", C.c_int), ("taper_strength", C.c_double),
                ("vti", C.c_int), ("gdamp_tmin", C.c_double), ("gdamp_tmax", C.c_double),
                ("gdamp_power", C.c_double)]


class OrcModelPar(C.Structure):
    _fields_ = [("vpmin", C.c_double), ("vpmax", C.c_double), ("vsmin", C.c_double), ("vsmax", C.c_double),
                ("rhomin", C.c_double), ("rhomax", C.c_double), ("deltamin", C.c_double),
                ("deltamax", C.c_double), ("epsmin", C.c_double), ("epsmax", C.c_double),
                ("soft_clip", C.c_int), ("courant", C.c_double), ("dt", C.c_double), ("sub", C.c_int),
                ("nt_data", C.c_int), ("dt_data", C.c_double), ("vmax", C.c_double), ("vmin", C.c_double),
                ("dt_prop", C.c_double), ("courant_out", C.c_double), ("nt_prop", C.c_int),
                ("sub_out", C.c_int)]


class OrcOp(C.Structure):
    _fields_ = [("g", OrcGrid), ("mp", OrcModelPar), ("ns", C.c_int), ("sxz", C.POINTER(C.c_double)),
                ("nrec", C.POINTER(C.c_int)), ("rxz", C.POINTER(C.c_double)), ("seismotype", C.c_int),
                ("resampling", C.c_int), ("sinc_hl", C.c_int), ("mt", C.c_int), ("gl", C.c_double),
                ("rdip", C.POINTER(C.c_double)), ("acoustic", C.c_int), ("stiffness", C.c_double)]


def grid_struct(p):
    """Build an OrcGrid from a dict-like problem description."""
    g = OrcGrid()
    for k, _ in OrcGrid._fields_:
        if k in p:
            setattr(g, k, p[k])
    return g


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class Oracle:
    """Numpy front end of the oracle restatement."""

    def __init__(self, bits=64):
        self.bits = bits
        path = os.path.join(HERE, "liboracle%d.so" % bits)
        if not os.path.exists(path):
            raise FileNotFoundError("%s missing: run `make -C oracle`" % path)
        self.lib = C.CDLL(path)
        self.dtype = np.float64 if bits == 64 else np.float32
        self.rt = C.c_double if bits == 64 else C.c_float
        assert self.lib.orc_real_size() == (8 if bits == 64 else 4)

    def arr(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def p(self, a):
        return _ptr(a, self.rt)

    # ---- operators --------------------------------------------------------
    def d2_table(self):
        out = np.zeros(384)
        self.lib.orc_d2_table(_ptr(out, C.c_double))
        return out.reshape(6, 8, 8)

    def Dx(self, u, d, p=None, add=None, a=1.0):
        u = self.arr(u)
        out = np.zeros_like(u) if add is None else self.arr(add).copy()
        pp = self.arr(p) if p is not None else None
        self.lib.orc_Dx(int(add is not None), self.p(u), self.p(out), self.p(pp), u.shape[0], u.shape[1],
                        C.c_double(d), C.c_double(a))
        return out

    def Dz(self, u, d, p=None, add=None, a=1.0):
        u = self.arr(u)
        out = np.zeros_like(u) if add is None else self.arr(add).copy()
        pp = self.arr(p) if p is not None else None
        self.lib.orc_Dz(int(add is not None), self.p(u), self.p(out), self.p(pp), u.shape[0], u.shape[1],
                        C.c_double(d), C.c_double(a))
        return out

    def Dxx_var(self, u, c, d):
        u, c = self.arr(u), self.arr(c)
        out = np.zeros_like(u)
        self.lib.orc_Dxx_var(0, self.p(u), self.p(out), self.p(c), u.shape[0], u.shape[1], C.c_double(d),
                             C.c_double(1.0))
        return out

    def Dzz_var(self, u, c, d):
        u, c = self.arr(u), self.arr(c)
        out = np.zeros_like(u)
        self.lib.orc_Dzz_var(0, self.p(u), self.p(out), self.p(c), u.shape[0], u.shape[1], C.c_double(d),
                             C.c_double(1.0))
        return out

    def applyH(self, u, d, axis_x, inv=False):
        u = self.arr(u).copy()
        self.lib.orc_applyH(int(axis_x), int(inv), self.p(u), u.shape[0], u.shape[1], C.c_double(d))
        return u

    def esat(self, side, in0, in1, in2, out, f1, f2, f3, dx, dz, dt):
        out = self.arr(out).copy()
        in0, in1, f1, f2 = self.arr(in0), self.arr(in1), self.arr(f1), self.arr(f2)
        in2 = self.arr(in2) if in2 is not None else None
        f3 = self.arr(f3) if f3 is not None else None
        self.lib.orc_esat(side, self.p(in0), self.p(in1), self.p(in2), self.p(out), self.p(f1), self.p(f2),
                          self.p(f3), out.shape[0], out.shape[1], C.c_double(dx), C.c_double(dz),
                          C.c_double(dt))
        return out

    def esat_dn(self, side, u, out, f, dx, dz):
        out = self.arr(out).copy()
        u, f = self.arr(u), self.arr(f)
        self.lib.orc_esat_dn(side, self.p(u), self.p(out), self.p(f), out.shape[0], out.shape[1],
                             C.c_double(dx), C.c_double(dz))
        return out

    def scale_boundaries(self, ux, uz, model, vti, dx, dz, dt, top, bottom, left, right):
        ux, uz, model = self.arr(ux).copy(), self.arr(uz).copy(), self.arr(model)
        self.lib.orc_scale_boundaries(self.p(ux), self.p(uz), self.p(model), int(vti), ux.shape[0],
                                      ux.shape[1], C.c_double(dx), C.c_double(dz), C.c_double(dt),
                                      int(top), int(bottom), int(left), int(right))
        return ux, uz

    def taper(self, u, ttop, tbot, tleft, tright, a):
        u = self.arr(u).copy()
        self.lib.orc_taper(self.p(u), u.shape[0], u.shape[1], ttop, tbot, tleft, tright, C.c_double(a))
        return u

    def delta_m3(self, nx, nz, dx, dz, ox, oz, locx, locz):
        locx = np.ascontiguousarray(locx, dtype=np.float64)
        locz = np.ascontiguousarray(locz, dtype=np.float64)
        n = len(locx)
        xi = np.zeros((n, 3), np.int32)
        zi = np.zeros((n, 3), np.int32)
        xw = np.zeros((n, 2, 3), self.dtype)
        zw = np.zeros((n, 2, 3), self.dtype)
        self.lib.orc_delta_m3(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(ox), C.c_double(oz), n,
                              _ptr(locx, C.c_double), _ptr(locz, C.c_double), _ptr(xi, C.c_int),
                              _ptr(zi, C.c_int), self.p(xw), self.p(zw))
        return xi, zi, xw, zw

    def rhs(self, g, model, curr, prev=None):
        model, curr = self.arr(model), self.arr(curr)
        prev = self.arr(prev) if prev is not None else None
        out = np.zeros_like(curr)
        rc = self.lib.orc_rhs(C.byref(g), self.p(model), self.p(curr), self.p(prev), self.p(out))
        assert rc == 0, rc
        return out

    def propagate(self, g, model, inj, inj_traces, ext=None, full=None, grad=None):
        """inj = (xi, zi, xw, zw) tables; inj_traces = [2][ninj][nt].  Returns receiver traces
        [2][next][nt] in forward mode (grad is None)."""
        model = self.arr(model)
        xi, zi, xw, zw = inj
        tr = self.arr(inj_traces)
        nt = g.nt
        assert tr.shape[-1] == nt
        if ext is not None:
            exi, ezi, exw, ezw = ext
            nr = exi.shape[0]
            out = np.zeros((2, nr, nt), self.dtype)
            ep = (_ptr(exi, C.c_int), _ptr(ezi, C.c_int), self.p(self.arr(exw)), self.p(self.arr(ezw)))
            outx, outz = self.p(out[0]), self.p(out[1])
        else:
            nr, out, ep, outx, outz = 0, None, (None, None, None, None), None, None
        rc = self.lib.orc_propagate(C.byref(g), self.p(model), xi.shape[0], _ptr(xi, C.c_int), _ptr(zi, C.c_int),
                                    self.p(self.arr(xw)), self.p(self.arr(zw)), self.p(tr[0]), self.p(tr[1]),
                                    nr, ep[0], ep[1], ep[2], ep[3], outx, outz,
                                    self.p(full), self.p(grad))
        assert rc == 0, "orc_propagate rc=%d" % rc
        return out

    # ---- acoustic operator (nl_we_op_a) ----------------------------------
    def asat(self, side, kind, in0, in1, out, K, ri, dx, dz, dt, a):
        """out + the acoustic SAT of one side (kind 1 Dirichlet, 2 absorbing, 3 Neumann)."""
        out = self.arr(out).copy()
        in0, K, ri = self.arr(in0), self.arr(K), self.arr(ri)
        in1 = self.arr(in1) if in1 is not None else in0
        self.lib.orc_asat(side, kind, self.p(in0), self.p(in1), self.p(out), self.p(K), self.p(ri), out.shape[0],
                          out.shape[1], C.c_double(dx), C.c_double(dz), C.c_double(dt), C.c_double(a))
        return out

    def asat_scale(self, u, K, ri, dx, dz, dt, top, bottom, left, right):
        u = self.arr(u).copy()
        self.lib.orc_asat_scale(self.p(u), self.p(self.arr(K)), self.p(self.arr(ri)), u.shape[0], u.shape[1],
                                C.c_double(dx), C.c_double(dz), C.c_double(dt), top, bottom, left, right)
        return u

    def propagate_acoustic(self, g, model, inj, inj_traces, ext=None, full=None, grad=None, alpha=0.2508560249):
        """Acoustic propagation, model (K, 1/rho) [2][nx][nz]; inj = (xi, zi, xw, zw) tables,
        inj_traces [ninj][nt].  Returns receiver traces [next][nt] in forward mode.  full:
        [nsnap][nx][nz] snapshots (written forward, read with grad [2][nx][nz])."""
        model = self.arr(model)
        xi, zi, xw, zw = inj
        tr = self.arr(inj_traces)
        nt = g.nt
        assert tr.shape[-1] == nt
        if ext is not None:
            exi, ezi, exw, ezw = ext
            nr = exi.shape[0]
            out = np.zeros((nr, nt), self.dtype)
            ep = (_ptr(exi, C.c_int), _ptr(ezi, C.c_int), self.p(self.arr(exw)), self.p(self.arr(ezw)))
        else:
            nr, out, ep = 0, None, (None, None, None, None)
        rc = self.lib.orc_propagate_acoustic(C.byref(g), self.p(model), C.c_double(alpha), xi.shape[0],
                                             _ptr(xi, C.c_int), _ptr(zi, C.c_int), self.p(self.arr(xw)),
                                             self.p(self.arr(zw)), self.p(tr), nr, ep[0], ep[1], ep[2], ep[3],
                                             self.p(out), self.p(full), self.p(grad))
        assert rc == 0, "orc_propagate_acoustic rc=%d" % rc
        return out

    # ---- model and traces ------------------------------------------------
    def analyze_model(self, mp, model, dx, dz):
        model = self.arr(model).copy()
        self.lib.orc_analyze_model(C.byref(mp), self.p(model), model.shape[0], model.shape[1],
                                   model.shape[2], C.c_double(dx), C.c_double(dz))
        return model

    def convert_model(self, model):
        m = self.arr(model).copy()
        self.lib.orc_convert_model(self.p(m), m.shape[0], C.c_long(m[0].size))
        return m

    def sinc(self, adjoint, ntm, dtm, ntd, dtd, hl, mod=None, dat=None):
        if not adjoint:
            mod = self.arr(mod)
            ntr = mod.size // ntm
            dat = np.zeros((ntr, ntd), self.dtype)
        else:
            dat = self.arr(dat)
            ntr = dat.size // ntd
            mod = np.zeros((ntr, ntm), self.dtype)
        self.lib.orc_sinc(int(adjoint), 0, ntr, ntm, C.c_double(dtm), ntd, C.c_double(dtd), hl, self.p(mod),
                          self.p(dat))
        return mod if adjoint else dat

    def Dt(self, adj, traces, dt):
        traces = self.arr(traces)
        nt = traces.shape[-1]
        out = np.zeros_like(traces)
        self.lib.orc_Dt(int(adj), traces.size // nt, nt, C.c_double(dt), self.p(traces), self.p(out))
        return out

    def Ht(self, inv, traces, dt):
        traces = self.arr(traces).copy()
        nt = traces.shape[-1]
        self.lib.orc_Ht(int(inv), traces.size // nt, nt, C.c_double(dt), self.p(traces))
        return traces

    # ---- operator level -------------------------------------------------
    def make_op(self, prob):
        """prob: dict with grid/bc/taper/vti keys, 'mp' dict, shots (sxz [ns,2]),
        receivers per shot (list of [n,2] arrays), seismotype, resampling, sinc_hl."""
        op = OrcOp()
        op.g = grid_struct(prob)
        mp = OrcModelPar()
        for k, v in prob["mp"].items():
            setattr(mp, k, v)
        op.mp = mp
        sxz = np.ascontiguousarray(prob["sxz"], dtype=np.float64)
        nrec = np.ascontiguousarray([len(r) for r in prob["rxz"]], dtype=np.int32)
        rxz = np.ascontiguousarray(np.concatenate(prob["rxz"], axis=0), dtype=np.float64)
        op.ns = len(sxz)
        op.sxz, op.nrec, op.rxz = _ptr(sxz, C.c_double), _ptr(nrec, C.c_int), _ptr(rxz, C.c_double)
        op.seismotype = prob.get("seismotype", 0)
        op.resampling = prob.get("resampling", 0)
        op.sinc_hl = prob.get("sinc_hl", 11)
        op.acoustic = int(prob.get("acoustic", 0))
        op.stiffness = float(prob.get("free_surface_stiffness", 1.05))
        op.mt = 0 if op.acoustic else int(prob.get("mt", 0))
        op.gl = 0.0 if op.acoustic else float(prob.get("gl", 0.0))
        rdip = np.ascontiguousarray(np.concatenate(prob["rdip"]) if "rdip" in prob else np.zeros(len(rxz)),
                                    dtype=np.float64)
        op.rdip = _ptr(rdip, C.c_double)
        op._keep = (sxz, nrec, rxz, rdip)
        return op

    def ddelta_m3(self, nx, nz, dx, dz, ox, oz, locx, locz):
        """ddelta_m3 footprints: (xind [n][3], zind [n][3], xw [n][2][6], zw [n][2][6])."""
        locx = np.ascontiguousarray(locx, np.float64)
        locz = np.ascontiguousarray(locz, np.float64)
        n = len(locx)
        xi, zi = np.zeros((n, 3), np.int32), np.zeros((n, 3), np.int32)
        xw, zw = np.zeros((n, 2, 6), self.dtype), np.zeros((n, 2, 6), self.dtype)
        self.lib.orc_ddelta_m3(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(ox), C.c_double(oz), n,
                               _ptr(locx, C.c_double), _ptr(locz, C.c_double), _ptr(xi, C.c_int), _ptr(zi, C.c_int),
                               self.p(xw), self.p(zw))
        return xi, zi, xw, zw

    def dipole_m3(self, nx, nz, dx, dz, ox, oz, locx, locz, dip, gl):
        """dipole_m3 footprints: (xind [n][6], zind [n][6], xw [n][2][6], zw [n][2][6])."""
        locx = np.ascontiguousarray(locx, np.float64)
        locz = np.ascontiguousarray(locz, np.float64)
        dip = np.ascontiguousarray(dip, np.float64)
        n = len(locx)
        xi, zi = np.zeros((n, 6), np.int32), np.zeros((n, 6), np.int32)
        xw, zw = np.zeros((n, 2, 6), self.dtype), np.zeros((n, 2, 6), self.dtype)
        self.lib.orc_dipole_m3(nx, nz, C.c_double(dx), C.c_double(dz), C.c_double(ox), C.c_double(oz), n,
                               _ptr(locx, C.c_double), _ptr(locz, C.c_double), _ptr(dip, C.c_double),
                               C.c_double(gl), _ptr(xi, C.c_int), _ptr(zi, C.c_int), self.p(xw), self.p(zw))
        return xi, zi, xw, zw

    def dipole_to_strain(self, adj, traces, dip, gl):
        """dipole_to_strain on [2][ntr][nt] (a copy)."""
        t = self.arr(traces).copy()
        dip = np.ascontiguousarray(dip, np.float64)
        self.lib.orc_dipole_to_strain(int(adj), self.p(t), _ptr(dip, C.c_double), t.shape[1], t.shape[2],
                                      C.c_double(gl))
        return t

    def apply_forward(self, op, model, allsrc, keep_full=False):
        model = self.arr(model)
        allsrc = self.arr(allsrc)
        nr = int(op._keep[1].sum())
        ncd, nfld = (1, 1) if op.acoustic else (2, 2)
        data = np.zeros((ncd, nr, op.mp.nt_data), self.dtype)
        full = None
        if keep_full:
            # size from a dry analysis of the model
            tmp = OrcModelPar.from_buffer_copy(op.mp)
            self.analyze_model(tmp, model, op.g.dx, op.g.dz)
            sub = tmp.sub_out
            nsnap = 1 + tmp.nt_prop // sub
            full = np.zeros((op.ns, nsnap, nfld, op.g.nx, op.g.nz), self.dtype)
        rc = self.lib.orc_apply_forward(C.byref(op), 0, self.p(model), self.p(allsrc), self.p(data), self.p(full))
        assert rc == 0, rc
        return data, full

    def apply_jacobianT(self, op, model0, data, full):
        model0 = self.arr(model0)
        data = self.arr(data)
        np_ = 2 if op.acoustic else (5 if op.g.vti else 3)
        grad = np.zeros((np_, op.g.nx, op.g.nz), self.dtype)
        rc = self.lib.orc_apply_jacobianT(C.byref(op), 0, self.p(grad), self.p(model0), self.p(data), self.p(full))
        assert rc == 0, rc
        return grad

    # ---- least-squares objective -----------------------------------------
    @staticmethod
    def nlls_functional(r, dt, dnorm=1.0):
        """getFunc of nlls_fwi / nlls_fwi_eco (optimization.hpp:375-385, 929-942):
        1/2 dt (sum r^2 - 1/2 sum over traces of the first and last samples squared) / dnorm,
        i.e. the trapezoid rule in time; accumulated in double precision."""
        tr = np.asarray(r, np.float64).reshape(-1, np.shape(r)[-1])
        ends = np.sum(tr[:, 0] ** 2 + tr[:, -1] ** 2)
        return 0.5 * dt * (np.sum(tr * tr) - 0.5 * ends) / dnorm

    @staticmethod
    def integral(u, dt, adjoint=False):
        """integral (operator.hpp:650-689): per trace y0 = dt/2 x0, y_t = y_t-1 + dt x_t,
        y_n-1 = y_n-2 + dt/2 x_n-1; the adjoint is the transposed sum."""
        u = np.array(u, np.float64)
        tr = u.reshape(-1, u.shape[-1])
        wts = np.full(tr.shape[1], dt)
        wts[0] = wts[-1] = 0.5 * dt
        if not adjoint:
            out = np.cumsum(tr * wts, axis=1)
            out[:, -1] = out[:, -2] + 0.5 * dt * tr[:, -1]
        else:
            out = np.cumsum((tr * dt)[:, ::-1], axis=1)[:, ::-1]
            out[:, -1] = 0.5 * dt * tr[:, -1]
            out[:, 0] *= 0.5
        return out.reshape(u.shape)

    @staticmethod
    def ttnormalize(u):
        """ttnormalize (misc.cpp:97-110): each trace over its L2 norm (1 below 1e-16).
        Returns (normalised, norms)."""
        tr = np.array(u, np.float64).reshape(-1, np.shape(u)[-1])
        n2 = np.sum(tr * tr, axis=1)
        n2[n2 <= 1e-16] = 1.0
        norms = np.sqrt(n2)
        return (tr / norms[:, None]).reshape(np.shape(u)), norms

    @staticmethod
    def conv1dnd(u, f, centered=True, adjoint=False):
        """conv1dnd (operator.hpp:1035-1057; operator.cpp:2092-2137): every trace convolved
        in time with the 1-D filter f, y[t] = sum_j f[j] x[t - j + sh], the terms with
        t - j + sh outside the trace left out; sh = floor((nf + 1)/2) - 1 when centred,
        else 0.  The adjoint scatters y[t] f[j] back to x[t - j + sh]."""
        u = np.asarray(u, np.float64)
        f = np.asarray(f, np.float64).ravel()
        tr = u.reshape(-1, u.shape[-1])
        nt, nf = tr.shape[1], f.size
        sh = (nf + 1) // 2 - 1 if centered else 0
        out = np.zeros_like(tr)
        for j in range(nf):
            k = sh - j  # y[t] takes x[t + k]
            lo, hi = max(0, -k), min(nt, nt - k)
            if lo >= hi:
                continue
            if not adjoint:
                out[:, lo:hi] += f[j] * tr[:, lo + k:hi + k]
            else:
                out[:, lo + k:hi + k] += f[j] * tr[:, lo:hi]
        return out.reshape(u.shape)

    @staticmethod
    def zero_phase(f):
        """zero_phase (operator.cpp:2197-2250): the filter padded to odd length n, its
        spectrum replaced by the amplitude spectrum times the linear phase
        exp(-i w (n-1)/2), w = 2 pi k/n; a symmetric filter of the same amplitude spectrum
        (centred by conv1dnd's shift)."""
        f = np.asarray(f, np.float64).ravel()
        if f.size % 2 == 0:
            f = np.concatenate([f, [0.0]])
        n = f.size
        k = np.arange(n // 2 + 1)
        X = np.abs(np.fft.rfft(f)) * np.exp(-1j * k * (2 * np.pi / n) * (n - 1) / 2)
        return np.fft.irfft(X, n)

    @staticmethod
    def minimum_phase(f, eps=1e-7):
        """minimum_phase (operator.cpp:2252-2320), the homomorphic construction: the filter
        padded to odd length n; the real cepstrum c = IFT(log(|FT f| + eps)); c folded onto
        its causal part (c_k doubled for 1 <= k <= n/2, zero above); the filter
        IFT(exp(FT c))."""
        f = np.asarray(f, np.float64).ravel()
        if f.size % 2 == 0:
            f = np.concatenate([f, [0.0]])
        n = f.size
        c = np.fft.irfft(np.log(np.abs(np.fft.rfft(f)) + eps) + 0j, n)
        c[1:n // 2 + 1] *= 2
        c[n // 2 + 1:] = 0
        return np.fft.irfft(np.exp(np.fft.rfft(c)), n)

    @classmethod
    def prepare_filter(cls, f, phase="default"):
        """The filter as nlls_fwi_eco's constructor prepares it (optimization.hpp:526-534):
        (filter, centred) for par.filter_phase "default" (as given), "zero" or "minimum"
        (not centred)."""
        if phase == "zero":
            return cls.zero_phase(f), True
        if phase == "minimum":
            return cls.minimum_phase(f), False
        return np.asarray(f, np.float64).ravel(), True

    def nlls_data_terms(self, u, d, dt, w=None, integrate=False, trace_normalize=False, filt=None, centered=True):
        """Data side of nlls_fwi / nlls_fwi_eco (optimization.hpp:330-455, 512-556, 660-905):
        with P(u) = ttnormalize(C W integral(u)) (each step optional; C the filter's
        conv1dnd), returns the residual r = P(u) - P(d), the adjoint source
        a = P'(u)^T Ht r, i.e. the gradient in u of 1/2 r^T Ht r, and P(d).
        P'^T = integral^T W C^T (I/|s| - s s^T/|s|^3) per trace, with s the filtered,
        weighted, integrated synthetics."""
        def process(v):
            v = np.asarray(v, np.float64)
            if integrate:
                v = self.integral(v, dt)
            if w is not None:
                v = v * w
            if filt is not None:
                v = self.conv1dnd(v, filt, centered)
            syn, norms = v, None
            if trace_normalize:
                v, norms = self.ttnormalize(v)
            return v, syn, norms
        pd = process(d)[0]
        pu, syn, norms = process(u)
        r = pu - pd
        a = self.Ht(False, np.array(r, self.dtype), dt)
        if trace_normalize:
            at = a.reshape(-1, a.shape[-1])
            st = syn.reshape(at.shape)
            v = np.sum(at * st, axis=1)
            a = (at / norms[:, None] - (v / norms ** 3)[:, None] * st).reshape(a.shape)
        if filt is not None:
            a = self.conv1dnd(a, filt, centered, adjoint=True)
        if w is not None:
            a = a * w
        if integrate:
            a = self.integral(a, dt, adjoint=True)
        return r, a, pd

    def nlls_evaluate(self, prob, model, allsrc, d, w=None, gmask=None, normalize=False, gradient=True,
                      integrate=False, trace_normalize=False, filt=None, filter_phase="default"):
        """nlls_fwi (optimization.hpp:309-455) or, with normalize, nlls_fwi_eco with
        normalize_obj_func (optimization.hpp:482-951); integrate and trace_normalize are
        par.integrate / par.normalize, filt and filter_phase nlls_fwi_eco's filter and
        par.filter_phase (prepared by prepare_filter): r, a from nlls_data_terms at u = F(m);
        f = nlls_functional(r, dt, dnorm); g = gmask J^T a / dnorm, with
        dnorm = dt |P d|^2 (1 below 1e-16) for nlls_fwi_eco and 1 for nlls_fwi.  The shot
        sum is one multi-shot apply_jacobianT(add=false), i.e. the per-shot calls of
        nlls_fwi_eco (optimization.hpp:890) summed.  Returns (f, g, r)."""
        dt = float(prob["mp"]["dt_data"])
        u, full = self.apply_forward(self.make_op(prob), model, allsrc, keep_full=gradient)
        centered = True
        if filt is not None:
            filt, centered = self.prepare_filter(filt, filter_phase)
        r, a, pd = self.nlls_data_terms(u, d, dt, w, integrate, trace_normalize, filt, centered)
        dnorm = 1.0
        if normalize:
            dnorm = float(np.sum(pd * pd)) * dt
            if dnorm < 1e-16:
                dnorm = 1.0
        f = self.nlls_functional(r, dt, dnorm)
        if not gradient:
            return f, None, r
        g = self.apply_jacobianT(self.make_op(prob), model, a, full) / dnorm
        if gmask is not None:
            g = g * gmask
        return f, g, r


class RefSBP:
    """The reference's own SBP templates and spatial_operators.cpp, compiled from source
    into _ref/libref_sbp.so (data_t = double)."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libref_sbp.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        assert self.lib.ref_data_t_size() == 8

    @staticmethod
    def p(a):
        return _ptr(a, C.c_double) if a is not None else None

    @staticmethod
    def arr(a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def Dzz_var(self, u, c, d):
        u, c = self.arr(u), self.arr(c)
        out = np.zeros_like(u)
        self.lib.ref_Dzz_var(0, self.p(u), self.p(out), self.p(c), u.shape[0], u.shape[1], C.c_double(d),
                             C.c_double(1.0))
        return out

    def Dxx_var(self, u, c, d):
        u, c = self.arr(u), self.arr(c)
        out = np.zeros_like(u)
        self.lib.ref_Dxx_var(0, self.p(u), self.p(out), self.p(c), u.shape[0], u.shape[1], C.c_double(d),
                             C.c_double(1.0))
        return out

    def asat(self, side, kind, in0, in1, out, K, ri, dx, dz, dt, a):
        out = self.arr(out).copy()
        in0, K, ri = self.arr(in0), self.arr(K), self.arr(ri)
        in1 = self.arr(in1) if in1 is not None else in0
        self.lib.ref_asat(side, kind, self.p(in0), self.p(in1), self.p(out), self.p(K), self.p(ri), out.shape[0],
                          out.shape[1], C.c_double(dx), C.c_double(dz), C.c_double(dt), C.c_double(a))
        return out

    def asat_scale(self, u, K, ri, dx, dz, dt, top, bottom, left, right):
        u = self.arr(u).copy()
        self.lib.ref_asat_scale(self.p(u), self.p(self.arr(K)), self.p(self.arr(ri)), u.shape[0], u.shape[1],
                                C.c_double(dx), C.c_double(dz), C.c_double(dt), top, bottom, left, right)
        return u

    def Dzz(self, u, d):
        u = self.arr(u)
        out = np.zeros_like(u)
        self.lib.ref_Dzz(0, self.p(u), self.p(out), u.shape[0], u.shape[1], C.c_double(d))
        return out

    def esat(self, side, in0, in1, in2, out, f1, f2, f3, dx, dz, dt):
        out = self.arr(out).copy()
        args = [self.arr(x) if x is not None else None for x in (in0, in1, in2, f1, f2, f3)]
        self.lib.ref_esat(side, self.p(args[0]), self.p(args[1]), self.p(args[2]), self.p(out), self.p(args[3]),
                          self.p(args[4]), self.p(args[5]), out.shape[0], out.shape[1], C.c_double(dx),
                          C.c_double(dz), C.c_double(dt))
        return out

    def esat_dn(self, side, u, out, f, dx, dz):
        out = self.arr(out).copy()
        u, f = self.arr(u), self.arr(f)
        self.lib.ref_esat_dn(side, self.p(u), self.p(out), self.p(f), out.shape[0], out.shape[1], C.c_double(dx),
                             C.c_double(dz))
        return out

    def scale_boundaries(self, ux, uz, model, vti, dx, dz, dt, top, bottom, left, right):
        ux, uz, model = self.arr(ux).copy(), self.arr(uz).copy(), self.arr(model)
        self.lib.ref_scale_boundaries(self.p(ux), self.p(uz), self.p(model), int(vti), ux.shape[0], ux.shape[1],
                                      C.c_double(dx), C.c_double(dz), C.c_double(dt), int(top), int(bottom),
                                      int(left), int(right))
        return ux, uz

    def applyH(self, u, d, axis_x, inv=False):
        u = self.arr(u)
        out = np.zeros_like(u)
        self.lib.ref_applyH(int(axis_x), int(inv), self.p(u), self.p(out), u.shape[0], u.shape[1], C.c_double(d))
        return out


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref_sbp.so"))
