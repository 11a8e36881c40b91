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
