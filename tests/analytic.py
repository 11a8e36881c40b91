"""Closed-form solutions used as physics pins (test helpers, no reference code).

Plane-strain elastic Green's tensor of a homogeneous full space, for
rho u_tt = mu lap u + (lambda + mu) grad div u + F(t) delta(x - xs) e_j:

    G_ij(w) = g_S(r) delta_ij / mu + d_i d_j (g_S - g_P)(r) / (rho w^2),

with g_c the 2-D Helmholtz Green's function, (lap + k_c^2) g_c = -delta, k_c = w / c, i.e.
g_c = -(i/4) H0^(2)(k_c r) for numpy's e^{+i w t} synthesis (outgoing waves).  Applying
the elastic operator to this G gives -delta_ij delta exactly (the g_S lap term and the
grad-div terms cancel with a = 1 / (rho w^2)).  For radial f(r):
d_i d_j f = f'' g_i g_j + f'/r (delta_ij - g_i g_j), g = (x - xs) / r.
"""
import numpy as np
from scipy.special import hankel2


def greens_traces(src, dt, vp, vs, rho, offsets, nt=None, pad=16):
    """Displacement traces [2][len(offsets)][nt] at receiver offsets (dx, dz) from a
    vertical point force src(t) (sampled at dt) in a homogeneous full space."""
    src = np.asarray(src, np.float64)
    nt = len(src) if nt is None else nt
    N = pad * len(src)
    F = np.fft.rfft(src, N) * dt
    w = 2 * np.pi * np.fft.rfftfreq(N, dt)[1:]
    mu = rho * vs ** 2

    def radial(k, r):
        x = k * r
        h0, h1 = hankel2(0, x), hankel2(1, x)
        return -0.25j * h0, 0.25j * k * h1, 0.25j * k ** 2 * (h0 - h1 / x)

    out = np.zeros((2, len(offsets), nt))
    for n, (ox, oz) in enumerate(offsets):
        r = float(np.hypot(ox, oz))
        gam = np.array([ox, oz]) / r
        s0, s1, s2 = radial(w / vs, r)
        p0, p1, p2 = radial(w / vp, r)
        for i in range(2):  # component i of the response to a force along z (j = 1)
            d = float(i == 1)
            dd = (s2 - p2) * gam[i] * gam[1] + (s1 - p1) / r * (d - gam[i] * gam[1])
            G = np.zeros(len(w) + 1, complex)
            G[1:] = s0 * d / mu + dd / (rho * w ** 2)
            out[i, n] = np.fft.irfft(G * F, N)[:nt] / dt
    return out


def homogeneous_case(vti=False, n=301, h=0.01, nt=1001, dt=0.001, f0=4.0):
    """A vertical 4 Hz Ricker force at the centre of a 3 km square (absorbing sides),
    receivers 0.25 km away in four directions (two off the grid): everything arrives
    before the first boundary reflection (1.4 s)."""
    vp, vs, rho = 2.0, 1.0, 2.0
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    fields = [lam, mu, rho] + ([lam, 0.0] if vti else [])  # VTI with eps = delta = 0: c13 = lambda
    model = np.stack([np.full((n, n), v) for v in fields])
    prob = dict(nx=n, nz=n, dx=h, dz=h, bc_top=2, bc_bottom=2, bc_left=2, bc_right=2, taper_top=20,
                taper_bottom=20, taper_left=20, taper_right=20, taper_strength=0.05, vti=int(vti),
                gdamp_tmin=0.0, gdamp_tmax=0.0, gdamp_power=1.0)
    xs = zs = (n - 1) * h / 2
    offsets = [(0.25, 0.0), (0.0, 0.25), (0.25 / np.sqrt(2), 0.25 / np.sqrt(2)), (-0.2, -0.15)]
    rxz = np.array([[xs + a, zs + b] for a, b in offsets])
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    wav = (1.0 - 2.0 * a) * np.exp(-a)
    src = np.zeros((2, 1, nt))
    src[1, 0] = wav
    ref = greens_traces(wav, dt, vp, vs, rho, offsets)
    return prob, model, np.array([[xs, zs]]), [rxz], src, nt, dt, ref


def rayleigh_speed(vp, vs):
    """Root of the Rayleigh equation (2 - c^2/vs^2)^2 = 4 sqrt(1 - c^2/vp^2) sqrt(1 - c^2/vs^2)."""
    from scipy.optimize import brentq
    f = lambda c: (2 - c ** 2 / vs ** 2) ** 2 - 4 * np.sqrt(1 - c ** 2 / vp ** 2) * np.sqrt(1 - c ** 2 / vs ** 2)
    return brentq(f, 0.5 * vs, 0.999 * vs)


def surface_case(n_x=401, n_z=301, h=0.01, nt=2901, dt=0.001, f0=4.0):
    """A vertical 4 Hz Ricker force on a free surface (absorbing elsewhere) and two surface
    receivers 1 and 2 km away; the Rayleigh pulse reaches both before any boundary
    reflection.  Returns prob, converted model, sxz, rxz, src, nt, dt, (vp, vs)."""
    vp, vs, rho = 2.0, 1.0, 2.0
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    model = np.stack([np.full((n_x, n_z), v) for v in (lam, mu, rho)])
    prob = dict(nx=n_x, nz=n_z, dx=h, dz=h, bc_top=1, bc_bottom=2, bc_left=2, bc_right=2, taper_top=0,
                taper_bottom=30, taper_left=30, taper_right=30, taper_strength=0.05, vti=0,
                gdamp_tmin=0.0, gdamp_tmax=0.0, gdamp_power=1.0)
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    src = np.zeros((2, 1, nt))
    src[1, 0] = (1.0 - 2.0 * a) * np.exp(-a)
    xs = 0.5
    rxz = np.array([[xs + 1.0, 0.0], [xs + 2.0, 0.0]])
    return prob, model, np.array([[xs, 0.0]]), [rxz], src, nt, dt, (vp, vs, f0)


def surface_wave_speed(uz, dt, c_expected, f0, offsets=(1.0, 2.0)):
    """Speed of the pulse between two surface traces: each trace windowed (Hann, 0.9 s)
    around the expected arrival, the lag of the cross-correlation peak refined by a
    parabola.  Returns (speed, peak amplitude ratio far/near)."""
    nt = uz.shape[-1]
    t = np.arange(nt) * dt

    def win(tr, x):
        m = np.abs(t - (x / c_expected + 1.2 / f0)) < 0.45
        out = np.zeros(nt)
        out[m] = tr[m] * np.hanning(m.sum())
        return out
    a1, a2 = win(uz[0], offsets[0]), win(uz[1], offsets[1])
    cc = np.correlate(a2, a1, "full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    lag = (k - (nt - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)) * dt
    return (offsets[1] - offsets[0]) / lag, np.abs(a2).max() / np.abs(a1).max()


def christoffel_vti(c11, c13, c33, c55, rho, theta):
    """Phase velocities (qP, qSV) of a plane wave in a VTI medium whose normal makes the
    angle theta (radians) with the symmetry axis z: rho v^2 are the eigenvalues of the
    Christoffel matrix [[c11 s^2 + c55 c^2, (c13 + c55) s c], [(c13 + c55) s c, c55 s^2 + c33 c^2]]
    with s = sin(theta), c = cos(theta)."""
    s, c = np.sin(theta), np.cos(theta)
    G = np.array([[c11 * s * s + c55 * c * c, (c13 + c55) * s * c], [(c13 + c55) * s * c, c55 * s * s + c33 * c * c]])
    ev = np.linalg.eigvalsh(G)
    return np.sqrt(ev[1] / rho), np.sqrt(ev[0] / rho)


VTI_PLANE = dict(vp0=2.0, vs0=1.0, rho=2.0, eps=0.2, delta=0.1)


def plane_wave_case(theta_deg, mode, n=481, h=0.005, nt=1000, dt=0.0008, f0=10.0, d=(0.2, 0.5), vti=VTI_PLANE):
    """A plane wave from a line of point forces through the centre of a homogeneous VTI
    square (absorbing sides), its normal n at theta_deg from the z axis (towards +x),
    forces along n (mode "P") or along the line (mode "SV"); two receivers on the normal
    d km from the line.  Returns prob, converted model (lambda, mu, rho, c13, eps), source
    positions [k][2], receivers [2][2], source traces [2][k][nt], the polarisation vector
    to project on, and the expected phase velocity of the mode (Christoffel)."""
    p = vti
    lam, mu = p["rho"] * (p["vp0"] ** 2 - 2 * p["vs0"] ** 2), p["rho"] * p["vs0"] ** 2
    c13 = np.sqrt(2 * (lam + 2 * mu) * (lam + mu) * p["delta"] + (lam + mu) ** 2) - mu
    c11, c33, c55 = (1 + 2 * p["eps"]) * (lam + 2 * mu), lam + 2 * mu, mu
    th = np.radians(theta_deg)
    vP, vS = christoffel_vti(c11, c13, c33, c55, p["rho"], th)
    model = np.stack([np.full((n, n), v) for v in (lam, mu, p["rho"], c13, p["eps"])])
    taper = 40
    prob = dict(nx=n, nz=n, dx=h, dz=h, bc_top=2, bc_bottom=2, bc_left=2, bc_right=2, taper_top=taper,
                taper_bottom=taper, taper_left=taper, taper_right=taper, taper_strength=0.05, vti=1,
                gdamp_tmin=0.0, gdamp_tmax=0.0, gdamp_power=1.0)
    L = (n - 1) * h
    c = np.array([L / 2, L / 2])
    nv = np.array([np.sin(th), np.cos(th)])
    tv = np.array([np.cos(th), -np.sin(th)])
    lo, hi = 3 * h, L - 3 * h
    pts = [c + s * tv for s in np.arange(-L, L, h)]
    pts = np.array([q for q in pts if lo <= q[0] <= hi and lo <= q[1] <= hi])
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    wav = (1.0 - 2.0 * a) * np.exp(-a)
    fdir = nv if mode == "P" else tv
    src = np.stack([np.tile(wav * fdir[0], (len(pts), 1)), np.tile(wav * fdir[1], (len(pts), 1))])
    rxz = np.array([c + di * nv for di in d])
    v = vP if mode == "P" else vS
    return prob, model, pts, rxz, src, fdir, v


def plane_wave_speed(traces, pol, dt, v_expected, f0=10.0, d=(0.2, 0.5)):
    """Phase velocity between the two receivers of plane_wave_case: the displacement
    projected on pol, each trace windowed (Hann, +/- 1.1/f0) around the expected arrival
    of the mode, the lag of the cross-correlation peak refined by a parabola."""
    u = pol[0] * traces[0] + pol[1] * traces[1]
    nt = u.shape[-1]
    t = np.arange(nt) * dt

    def win(tr, x):
        m = np.abs(t - (x / v_expected + 1.2 / f0)) < 1.1 / f0
        out = np.zeros(nt)
        out[m] = tr[m] * np.hanning(m.sum())
        return out
    a1, a2 = win(u[0], d[0]), win(u[1], d[1])
    cc = np.correlate(a2, a1, "full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    lag = (k - (nt - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)) * dt
    return (d[1] - d[0]) / lag


def acoustic_greens_traces(src, dt, v, rho, offsets, nt=None, pad=16):
    """Pressure traces [len(offsets)][nt] at receiver offsets (dx, dz) from a point source
    s(t) of the acoustic operator, u_tt = K (div(grad u / rho) + s(t) delta(x - xs)), in a
    homogeneous full space: u = K g_v * s / v^2 = rho g_v * s with the 2-D Helmholtz
    Green's function g_v of the module docstring."""
    src = np.asarray(src, np.float64)
    nt = len(src) if nt is None else nt
    N = pad * len(src)
    F = np.fft.rfft(src, N) * dt
    w = 2 * np.pi * np.fft.rfftfreq(N, dt)[1:]
    out = np.zeros((len(offsets), nt))
    for n, (ox, oz) in enumerate(offsets):
        r = float(np.hypot(ox, oz))
        G = np.zeros(len(w) + 1, complex)
        G[1:] = rho * (-0.25j) * hankel2(0, w / v * r)
        out[n] = np.fft.irfft(G * F, N)[:nt] / dt
    return out
