"""Synthetic workloads: elastic models, Ricker wavelets and acquisition geometry.

Units follow the reference operators: km, km/s, g/cc and seconds (the reference
hard-clips models to vpmax=8 km/s, rhomax=8 g/cc in analyzeModel,
we_op.cpp:334-338, so SI inputs must be converted first).

A *problem* is a plain dict consumed by both the engine (fwi2d_amd.op) and the
CPU reference restatement used by the tests:

    nx, nz, dx, dz, ox, oz          grid (axis 1 = z fastest, axis 2 = x)
    bc_top/bottom/left/right        1 free surface, 2 absorbing
    taper_top/bottom/left/right, taper_strength
    vti                             0 isotropic (vp, vs, rho), 1 VTI (vp, vs, rho, delta, eps)
    mp                              model bounds and time-axis requests (param.hpp:10-41)
    sxz [ns,2]                      source positions (x, z)
    rxz list of [n_s,2]             receiver positions per shot (uneven counts allowed)
    seismotype                      0 displacement, 1 particle velocity
    resampling, sinc_hl             0 sinc (default, half length 11), 1 linear
"""
import numpy as np

# param.hpp:10-41 defaults, as a dict of the fields the path uses
DEFAULT_MP = dict(vpmin=0.2, vpmax=8.0, vsmin=0.1, vsmax=5.0, rhomin=0.1, rhomax=8.0,
                  deltamin=-0.5, deltamax=1.0, epsmin=0.0, epsmax=1.0, soft_clip=0,
                  courant=0.6, dt=-1.0, sub=-1)


def ricker(nt, dt, f0, t0=None):
    """Ricker wavelet w(t) = (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - t0,
    t0 = 1.2/f by default."""
    t = np.arange(nt) * dt
    t0 = 1.2 / f0 if t0 is None else t0
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _smooth_field(shape, sigma_cells, rng):
    from scipy.ndimage import gaussian_filter
    p = gaussian_filter(rng.standard_normal(shape), sigma=sigma_cells, mode="reflect")
    p -= p.mean()
    s = p.std()
    return p / s if s > 0 else p


def gradient_model(nx, nz, dx, dz, vp_top=1.5, vp_bot=3.5, pert=0.05, corr=0.2, seed=0):
    """vp(z) linear from vp_top to vp_bot km/s, times (1 + pert * p) where p is a seeded
    N(0,1) field Gaussian-filtered with sigma = corr/(h*sqrt(2)) and renormalised;
    vs = vp/1.8; rho = 0.31 * (1000 vp)^0.25 g/cc (Gardner).  Returns [3][nx][nz]."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * dz
    zmax = max(z[-1], 1e-12)
    vp = np.tile(vp_top + (vp_bot - vp_top) * z / zmax, (nx, 1))
    if pert:
        sig = (corr / (dx * np.sqrt(2.0)), corr / (dz * np.sqrt(2.0)))
        vp = vp * (1.0 + pert * _smooth_field((nx, nz), sig, rng))
    vs = vp / 1.8
    rho = 0.31 * (1000.0 * vp) ** 0.25
    return np.stack([vp, vs, rho])


def add_ellipses(model, dx, dz, n=2, amp=0.10, seed=1):
    """Add n seeded ellipsoidal anomalies of +/- amp on vp and vs (rho via Gardner)."""
    rng = np.random.default_rng(seed)
    m = model.copy()
    nx, nz = m.shape[1:]
    X, Z = np.meshgrid(np.arange(nx) * dx, np.arange(nz) * dz, indexing="ij")
    for k in range(n):
        cx = rng.uniform(0.25, 0.75) * (nx - 1) * dx
        cz = rng.uniform(0.35, 0.75) * (nz - 1) * dz
        ax = rng.uniform(0.05, 0.12) * (nx - 1) * dx
        az = rng.uniform(0.05, 0.12) * (nz - 1) * dz
        sgn = 1.0 if k % 2 == 0 else -1.0
        mask = ((X - cx) / ax) ** 2 + ((Z - cz) / az) ** 2 <= 1.0
        m[0][mask] *= 1.0 + sgn * amp
        m[1][mask] *= 1.0 + sgn * amp
    m[2] = 0.31 * (1000.0 * m[0]) ** 0.25
    return m


def smooth_model(model, dx, dz, length):
    """Gaussian smoothing of every parameter with sigma = length (km)."""
    from scipy.ndimage import gaussian_filter
    out = np.empty_like(model)
    for i in range(model.shape[0]):
        out[i] = gaussian_filter(model[i], sigma=(length / dx, length / dz), mode="nearest")
    return out


def vti_fields(nx, nz, dx, dz, eps_range=(0.0, 0.25), delta_range=(-0.05, 0.15), corr=0.3, seed=3):
    """Seeded smooth Thomsen delta and epsilon fields spanning the given ranges."""
    rng = np.random.default_rng(seed)
    sig = (corr / (dx * np.sqrt(2.0)), corr / (dz * np.sqrt(2.0)))
    out = []
    for lo, hi in (delta_range, eps_range):
        p = _smooth_field((nx, nz), sig, rng)
        p = (p - p.min()) / max(p.max() - p.min(), 1e-12)
        out.append(lo + (hi - lo) * p)
    return np.stack(out)  # (delta, eps)


def base_problem(nx, nz, dx, dz, nt_data, dt_data, sxz, rxz, vti=0, taper=30, taper_strength=0.05,
                 bc=(1, 2, 2, 2), seismotype=1, dt=-1.0, sub=-1, courant=0.6):
    mp = dict(DEFAULT_MP)
    mp.update(nt_data=int(nt_data), dt_data=float(dt_data), dt=float(dt), sub=int(sub), courant=courant)
    if vti:
        mp.update(epsmin=0.0, epsmax=1.0)
    return dict(nx=int(nx), nz=int(nz), dx=float(dx), dz=float(dz), ox=0.0, oz=0.0,
                bc_top=bc[0], bc_bottom=bc[1], bc_left=bc[2], bc_right=bc[3],
                taper_top=0 if bc[0] == 1 else taper, taper_bottom=taper if bc[1] == 2 else 0,
                taper_left=taper if bc[2] == 2 else 0, taper_right=taper if bc[3] == 2 else 0,
                taper_strength=float(taper_strength), vti=int(vti), mp=mp,
                sxz=np.asarray(sxz, dtype=np.float64).reshape(-1, 2),
                rxz=[np.asarray(r, dtype=np.float64).reshape(-1, 2) for r in rxz],
                seismotype=int(seismotype), resampling=0, sinc_hl=11,
                gdamp_tmin=0.0, gdamp_tmax=0.0, gdamp_power=1.0)


def vertical_force(wavelet, ns, fangle=1.5708):
    """analyzeWavelet for a vector force (we_op.cpp:196-217): [2][ns][nt] with
    fx = w cos(fangle), fz = w sin(fangle)."""
    w = np.asarray(wavelet, dtype=np.float64)
    return np.stack([np.tile(w * np.cos(fangle), (ns, 1)), np.tile(w * np.sin(fangle), (ns, 1))])


# ---------------------------------------------------------------------------
# The five benchmark configurations (sizes in cells; h in km)
# ---------------------------------------------------------------------------
def config0(seed=0):
    """Plain isotropic forward modelling, 1 shot: 201 x 601, h = 10 m, dt = 1 ms, nt = 2000,
    15 Hz Ricker at (3.0 km, 20 m), 150 surface receivers every 40 m."""
    nz, nx, h = 201, 601, 0.01
    nt, dt = 2000, 0.001
    model = gradient_model(nx, nz, h, h, seed=seed)
    rx = np.stack([np.arange(150) * 0.04, np.zeros(150)], axis=1)
    prob = base_problem(nx, nz, h, h, nt, dt, [[3.0, 0.02]], [rx], dt=dt, sub=0)
    src = vertical_force(ricker(nt, dt, 15.0), 1)
    return prob, model, src


def config1(seed=0):
    """Single-shot isotropic FWI gradient: 300 x 1000, h = 10 m, nt = 3000, dt = 1 ms,
    300 receivers, true = gradient model + 2 ellipses, start = 300 m smoothed."""
    nz, nx, h = 300, 1000, 0.01
    nt, dt = 3000, 0.001
    true = add_ellipses(gradient_model(nx, nz, h, h, seed=seed), h, h, seed=seed + 1)
    start = smooth_model(true, h, h, 0.3)
    rx = np.stack([np.linspace(0.1, 9.8, 300), np.full(300, 0.02)], axis=1)
    prob = base_problem(nx, nz, h, h, nt, dt, [[5.0, 0.02]], [rx], dt=dt, sub=-1)
    src = vertical_force(ricker(nt, dt, 10.0), 1)
    return prob, true, start, src


def config2_geometry(nshots=64, nx=1700, h=0.01, seed=2):
    """Seeded source positions U(0.3, 16.7) km, split-spread receivers every 40 m with
    shot-dependent aperture giving 100-400 receivers, peak frequencies U(8, 20) Hz."""
    rng = np.random.default_rng(seed)
    xmax = (nx - 1) * h
    sx = rng.uniform(0.3, min(16.7, xmax - 0.3), nshots)
    freqs = rng.uniform(8.0, 20.0, nshots)
    nrec = rng.integers(100, 401, nshots)
    rxz = []
    for s in range(nshots):
        n = int(nrec[s])
        half = (n - 1) * 0.04 / 2
        lo = min(max(sx[s] - half, 0.0), max(xmax - (n - 1) * 0.04, 0.0))
        x = lo + np.arange(n) * 0.04
        x = x[x <= xmax]
        rxz.append(np.stack([x, np.full(len(x), 0.02)], axis=1))
    sxz = np.stack([sx, np.full(nshots, 0.02)], axis=1)
    return sxz, rxz, freqs


def config2(nshots=64, seed=0):
    """Batched gradients: 64 shots, 500 x 1700, h = 10 m, nt = 4000 (dt = 1 ms)."""
    nz, nx, h = 500, 1700, 0.01
    nt, dt = 4000, 0.001
    model = gradient_model(nx, nz, h, h, seed=seed)
    sxz, rxz, freqs = config2_geometry(nshots, nx, h)
    prob = base_problem(nx, nz, h, h, nt, dt, sxz, rxz, dt=dt, sub=-1)
    wav = np.stack([ricker(nt, dt, f) for f in freqs])
    src = np.stack([wav * np.cos(1.5708), wav * np.sin(1.5708)])
    return prob, model, src


def config3(nshots=16, seed=0):
    """VTI gradient: 16 shots, 1000 x 3000, h = 5 m, nt = 6000, dt = 0.5 ms."""
    nz, nx, h = 1000, 3000, 0.005
    nt, dt = 6000, 0.0005
    m = gradient_model(nx, nz, h, h, seed=seed)
    de = vti_fields(nx, nz, h, h)
    model = np.concatenate([m, de], axis=0)
    sxz, rxz, freqs = config2_geometry(nshots, nx, h, seed=3)
    prob = base_problem(nx, nz, h, h, nt, dt, sxz, rxz, vti=1, dt=dt, sub=-1)
    wav = np.stack([ricker(nt, dt, f) for f in freqs])
    src = np.stack([wav * np.cos(1.5708), wav * np.sin(1.5708)])
    return prob, model, src


def notebook_elastic_fwi():
    """The published elastic FWI example of the reference (examples/elastic_fwi.ipynb
    cells 3, 5, 8 and 11), as the reference programs see it.

    - Wavelet (cell 3): 1001 samples of 4 ms, w = (1 - (tau/sig)^2) exp(-(tau/sig)^2 / 2),
      tau = t - 50 dt, sig = sqrt(2)/(2 pi f0), f0 = 10 Hz, stored as float32.
    - Models (cell 5): 801 x 201 cells of 10 m; two layers (vp, vs, rho) = (2.5, 1.5, 1.0)
      above z index 100 and (4.0, 2.0, 2.0) below; the true model adds (3.5, 1.75, 1.75)
      in x indices 375-425, z indices 70-79.
    - Geometry (cells 8, 11): ns=10 sx0=2 sz0=0.1 sxinc=0.4 nr=401 rx0=2 rz0=0.1
      rxinc=0.01, built as readParameters does (src/cpp/param.cpp:226-244) in the
      reference's single-precision data_t: x = x0 + float(i) * inc in float32.
    - Operator: bc_top=1, other sides 2 with 50-cell tapers of strength 0.08, dt=-1
      (courant 0.6), seismotype=1, fangle=1.5708 (analyzeWavelet, src/cpp/we_op.cpp:
      196-217, scaling in data_t), sub=-1 (analyzeNLInversion, src/cpp/param.cpp:253).

    Returns (prob, true_model, start_model, allsrc [2][10][1001] float32)."""
    f32 = np.float32
    nt, dt, f0, shift = 1001, 0.004, 10.0, 50
    t = np.linspace(0, (nt - 1) * dt, nt)
    sig = np.sqrt(2) / (2 * f0 * np.pi)
    w = ((1.0 - ((t - shift * dt) / sig) ** 2) * np.exp(-0.5 * ((t - shift * dt) / sig) ** 2)).astype(f32)
    nz, nx, h = 201, 801, 0.01
    m0 = np.zeros((3, nx, nz))
    m0[0, :, :100], m0[0, :, 100:] = 2.5, 4.0
    m0[1, :, :100], m0[1, :, 100:] = 1.5, 2.0
    m0[2, :, :100], m0[2, :, 100:] = 1.0, 2.0
    true = m0.copy()
    true[0, 375:426, 70:80], true[1, 375:426, 70:80], true[2, 375:426, 70:80] = 3.5, 1.75, 1.75
    ns, nr = 10, 401
    sx = f32(2) + np.arange(ns, dtype=f32) * f32(0.4)
    sxz = np.stack([sx, np.full(ns, f32(0.1), f32)], 1).astype(np.float64)
    rx = f32(2) + np.arange(nr, dtype=f32) * f32(0.01)
    rxz = np.stack([rx, np.full(nr, f32(0.1), f32)], 1).astype(np.float64)
    prob = base_problem(nx, nz, float(f32(h)), float(f32(h)), nt, float(f32(dt)), sxz, [rxz] * ns,
                        taper=50, taper_strength=float(f32(0.08)), dt=-1.0, sub=-1)
    fa = f32(1.5708)
    allsrc = np.stack([np.tile(w * f32(np.cos(fa)), (ns, 1)), np.tile(w * f32(np.sin(fa)), (ns, 1))])
    return prob, true.astype(f32), m0.astype(f32), allsrc.astype(f32)


def notebook_acoustic_fwi():
    """The published acoustic FWI example of the reference (examples/acoustic_fwi.ipynb
    cells 3, 5, 8 and 11; examples/par/acoustic_fwi.p), as the reference programs see it.

    - Wavelet (cell 3): the elastic example's 10 Hz Ricker, 1001 samples of 4 ms, shifted
      50 samples, float32; duplicated once per shot as the acoustic source (analyzeWavelet
      with nmodels = 2, src/cpp/we_op.cpp:178-194).
    - Models (cell 5): 801 x 201 cells of 10 m; (vp, rho) = (1.5, 1.0) above z index 100
      and (2.5, 2.0) below; the true model has (2.0, 1.75) in x indices 375-425, z indices
      70-79.
    - Geometry (cell 8, acoustic_fwi.p): ns=10 sx0=2 sz0=0.1 sxinc=0.4 nr=401 rx0=2 rz0=0.1
      rxinc=0.01 in float32 as readParameters builds it (src/cpp/param.cpp:226-244).
    - Operator: bc_top=1, other sides 2 with 50-cell tapers of strength 0.08, dt=-1
      (courant 0.6), seismotype=0 (pressure), sub=-1, free_surface_stiffness 1.05 (the
      default, src/include/param.hpp:35); the inversion's bounds vp in [1, 3] and rho in
      [1, 2.5] (acoustic_fwi.p).

    Returns (prob, true_model, start_model, allsrc [1][10][1001] float32)."""
    f32 = np.float32
    nt, dt, f0, shift = 1001, 0.004, 10.0, 50
    t = np.linspace(0, (nt - 1) * dt, nt)
    sig = np.sqrt(2) / (2 * f0 * np.pi)
    w = ((1.0 - ((t - shift * dt) / sig) ** 2) * np.exp(-0.5 * ((t - shift * dt) / sig) ** 2)).astype(f32)
    nz, nx, h = 201, 801, 0.01
    m0 = np.zeros((2, nx, nz))
    m0[0, :, :100], m0[0, :, 100:] = 1.5, 2.5
    m0[1, :, :100], m0[1, :, 100:] = 1.0, 2.0
    true = m0.copy()
    true[0, 375:426, 70:80], true[1, 375:426, 70:80] = 2.0, 1.75
    ns, nr = 10, 401
    sx = f32(2) + np.arange(ns, dtype=f32) * f32(0.4)
    sxz = np.stack([sx, np.full(ns, f32(0.1), f32)], 1).astype(np.float64)
    rx = f32(2) + np.arange(nr, dtype=f32) * f32(0.01)
    rxz = np.stack([rx, np.full(nr, f32(0.1), f32)], 1).astype(np.float64)
    prob = base_problem(nx, nz, float(f32(h)), float(f32(h)), nt, float(f32(dt)), sxz, [rxz] * ns,
                        taper=50, taper_strength=float(f32(0.08)), seismotype=0, dt=-1.0, sub=-1)
    prob["acoustic"] = 1
    prob["free_surface_stiffness"] = 1.05
    prob["mp"].update(vpmin=1.0, vpmax=3.0, rhomin=1.0, rhomax=2.5)
    allsrc = np.tile(w, (ns, 1))[None]
    return prob, true.astype(f32), m0.astype(f32), allsrc.astype(f32)


def config4(nshots=48, f0=5.0, seed=0):
    """End-to-end FWI: 48 shots, 600 x 2000, h = 10 m, nt = 4000 (dt = 1 ms); true = depth
    gradient + seeded lens anomalies, start = the true model smoothed over 500 m; Ricker
    sources of peak frequency f0 (one band of the multiscale schedule) on the 1 ms data
    axis.  Physical models (vp, vs, rho) and operator-level problem."""
    nz, nx, h = 600, 2000, 0.01
    nt, dt = 4000, 0.001
    true = add_ellipses(gradient_model(nx, nz, h, h, seed=seed), h, h, n=4, amp=0.08, seed=seed + 7)
    start = smooth_model(true, h, h, 0.5)
    sxz, rxz, _ = config2_geometry(nshots, nx, h, seed=seed + 4)
    # propagation step and snapshot interval from the model (courant 0.6), as in FWI runs
    # where the line search may try models with larger velocities
    prob = base_problem(nx, nz, h, h, nt, dt, sxz, rxz, dt=-1.0, sub=-1)
    src = vertical_force(ricker(nt, dt, f0), nshots)
    return prob, true, start, src
