"""The reference's published elastic FWI run, reproduced on the HIP path.

examples/elastic_fwi.ipynb (cells 3, 5, 8, 11) generates data with WE_MODELING.x on a
two-layer model with an intrusion and inverts them with FWI2D.x from the two-layer
model: nlls_fwi_eco without normalisation (src/main/fwi2d.cpp:339-341), l-BFGS with
memory 5 and the regular Wolfe line search at the defaults of src/include/param.hpp:62-77
(src/main/fwi2d.cpp:343-350).  Its printed output (cell 11) is the reference-held pin of
the whole forward and adjoint composition: the iteration-0 functional pins the
propagation (taper composition, SAT terms, sinc resampling, Dt, trapezoid misfit), the
first step length a0/|g0| the gradient (imaging condition with its lag, H quadrature,
chain rule), and the rest of the trace the solver on top.

The run here is the same through the C ABI: WeOp.apply_forward of the true model for the
data, the Nlls objective and fwi2d_lbfgs with the reference defaults.
"""
import numpy as np
import pytest

from fwi2d_amd import workloads as W

pytestmark = pytest.mark.gpu

# elastic_fwi.ipynb cell 11 output: functional after each iteration, and every trial
# (iteration, trial, step length, function value, gradient norm or None for "Last
# gradient norm")
PUBLISHED_FUNC = [16.598963, 11.779595, 10.589397, 7.140383, 6.211187, 5.269199, 4.540046, 4.353827, 3.699440,
                  3.358502, 2.943727]
PUBLISHED_TRIALS = [
    (0, 1, 0.125114, 11.779595, 4.895442),
    (1, 1, 2.666843, 10.589397, 11.487027),
    (2, 1, 1.489211, 7.140383, 4.773777),
    (3, 1, 8.323133, 17.040951, None),
    (3, 2, 1.858919, 6.211187, 2.811370),
    (4, 1, 3.588009, 5.269199, 2.431798),
    (5, 1, 1.525701, 4.540046, 2.579002),
    (6, 1, 1.611089, 4.353827, 4.268553),
    (7, 1, 1.573950, 3.699440, 1.511143),
    (8, 1, 6.466815, 4.293114, None),
    (8, 2, 2.426447, 3.358502, 1.615726),
    (9, 1, 2.385426, 2.943727, 1.087129),
]
PUBLISHED_FEVAL, PUBLISHED_GEVAL = 13, 11
# OpenMP threads assumed for the reference run's float reductions (not printed by it)
REF_THREADS = 16


def rel(a, b):
    return abs(a - b) / abs(b)


def h_weights(n, d):
    """Diagonal of the SBP quadrature H along one axis (applyHz/applyHx,
    src/cpp/spatial_operators.cpp:6-98): d * (17/48, 59/48, 43/48, 49/48, 1, ..., mirrored)."""
    w = np.ones(n)
    c = np.array([17.0, 59.0, 43.0, 49.0]) / 48
    w[:4] = c
    w[-4:] = c[::-1]
    return w * d


def cpu_add_true_accumulation(per_shot, model0, dx, dz):
    """The gradient the reference's CPU build returns when nlls_fwi_eco calls
    apply_jacobianT(add=true) shot after shot into one buffer
    (src/include/optimization.hpp:890; src/cpp/we_op.cpp:1080-1084, 1192-1236): the
    buffer is not cleared, so the shot's raw gradient is added to the gradient of the
    previous shots, H and the chain rule are applied to the sum, and the previous
    gradient is added again: G_s = g_s + T(G_s-1) + G_s-1, with g_s the shot's own
    gradient and T the H quadrature + chain rule at model0 after analyzeModel's hard clips
    (src/cpp/we_op.cpp:334-359, default bounds of src/include/param.hpp:43) and
    convert_model."""
    vp, vs, rho = [np.asarray(x, np.float64) for x in model0]
    vp, vs, rho = np.clip(vp, 0.2, 8.0), np.clip(vs, 0.1, 5.0), np.clip(rho, 0.1, 8.0)
    vs = np.where(vp < np.sqrt(2) * vs, vp / np.sqrt(2.0001), vs)
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    nx, nz = vp.shape
    H = np.outer(h_weights(nx, dx), h_weights(nz, dz))

    def T(G):
        a, b, c = G[0] * H, G[1] * H, G[2] * H
        return np.stack([2 * np.sqrt(rho * (lam + 2 * mu)) * a, 2 * np.sqrt(rho * mu) * (b - 2 * a),
                         lam / rho * a + mu / rho * b + c])

    G = np.zeros_like(per_shot[0], dtype=np.float64)
    for g in per_shot:
        G = g + T(G) + G
    return G


def reference_functional_fp32(r, dt, threads):
    """getFunc of nlls_fwi_eco as the reference's single-precision CPU build evaluates it
    (src/include/optimization.hpp:929-942): |r|^2 by vecReg::norm2 (src/include/vecReg.hpp:
    279-285), an OpenMP reduction whose threads each add their contiguous share of the
    squares one after the other in float (the build has no -ffast-math,
    CMakeLists.txt:12, so nothing is reassociated); the partial sums added in float; the
    end samples of every trace subtracted one trace after the other in float; then
    0.5 dt f / dnorm with dnorm = 1.  Float summation drops the low bits of every small
    square added to a large partial sum, so the result sits below the exact sum."""
    sq = np.asarray(r, np.float32).ravel()
    sq = sq * sq
    n = sq.size
    q, rem = divmod(n, threads)
    val, start = np.float32(0), 0
    for t in range(threads):
        m = q + (1 if t < rem else 0)
        if m:
            val = np.float32(val + np.add.accumulate(sq[start:start + m], dtype=np.float32)[-1])
        start += m
    tr = np.asarray(r, np.float32).reshape(-1, r.shape[-1])
    ends = (tr[:, 0] * tr[:, 0] + tr[:, -1] * tr[:, -1]).astype(np.float32)
    f = val
    for e in ends:
        f = np.float32(f - 0.5 * float(e))
    return float(np.float32(0.5 * float(np.float32(dt)) * float(f)))


@pytest.fixture(scope="module")
def run():
    import fwi2d_amd
    prob, true, start, allsrc = W.notebook_elastic_fwi()
    op = fwi2d_amd.WeOp(prob, allsrc)
    d = op.apply_forward(true)  # WE_MODELING.x
    axis = op.time_axis()
    obj = fwi2d_amd.Nlls(op, d)  # nlls_fwi_eco, normalize_obj_func=0
    f0, g0 = obj.evaluate(start)
    out = dict(prob=prob, start=start, op=op, obj=obj, axis=axis, f0=f0, g0=g0)
    nr = len(prob["rxz"][0])
    # per-shot gradients at the start model, for the CPU build's add=true accumulation
    r = obj.residual()
    out["r0"] = r.copy()
    dt = prob["mp"]["dt_data"]
    ht = np.full(r.shape[-1], dt)
    ht[0] = ht[-1] = 0.5 * dt
    a = r * ht
    per_shot = []
    for s in range(len(prob["sxz"])):
        a_s = np.zeros_like(a)
        a_s[:, s * nr:(s + 1) * nr] = a[:, s * nr:(s + 1) * nr]
        per_shot.append(op.apply_jacobianT(start, a_s).astype(np.float64))
    out["per_shot"] = per_shot
    # the same per-shot gradients without the reference's one-step imaging lag
    eng = op.engine_handle()
    lib = fwi2d_amd.load()
    assert lib.fwi2d_engine_set_option(eng, b"lag", 0) == 0
    per_shot_nolag = []
    for s in range(len(prob["sxz"])):
        a_s = np.zeros_like(a)
        a_s[:, s * nr:(s + 1) * nr] = a[:, s * nr:(s + 1) * nr]
        per_shot_nolag.append(op.apply_jacobianT(start, a_s).astype(np.float64))
    assert lib.fwi2d_engine_set_option(eng, b"lag", 1) == 0
    out["per_shot_nolag"] = per_shot_nolag
    m, hist, k, info = fwi2d_amd.lbfgs(obj, start, niter=10, report=True)
    out.update(hist=hist, iters=k, info=info, model=m)
    # the CPU build's run: nlls_fwi_eco's per-shot operators (optimization.hpp:618-669),
    # their gradients accumulated with apply_jacobianT(add=true) (optimization.hpp:890)
    ops, objs = [], []
    for s in range(len(prob["sxz"])):
        ps = dict(prob, sxz=prob["sxz"][s:s + 1], rxz=prob["rxz"][s:s + 1])
        ops.append(fwi2d_amd.WeOp(ps, allsrc[:, s:s + 1]))
        objs.append(fwi2d_amd.Nlls(ops[-1], d[:, s * nr:(s + 1) * nr]))

    def cpu_build_objective(model):
        fs, gs = zip(*(o.evaluate(model) for o in objs))
        G = cpu_add_true_accumulation([g.astype(np.float64) for g in gs], model, prob["dx"], prob["dz"])
        r = np.concatenate([o.residual() for o in objs], axis=1)
        return reference_functional_fp32(r, prob["mp"]["dt_data"], REF_THREADS), G.astype(np.float32)

    m, hist, k, info = fwi2d_amd.lbfgs(cpu_build_objective, start, niter=10, report=True)
    out.update(cpu_hist=hist, cpu_iters=k, cpu_info=info)
    return out


def test_time_axis(run):
    nt, dt, sub = run["axis"]
    assert nt == 3001 and sub == 3 and abs(dt - 0.004 / 3) < 1e-9, run["axis"]


def test_iteration0_functional(run):
    f0 = run["f0"]
    print("iteration-0 functional %.6f, published %.6f, rel %.2e" % (f0, PUBLISHED_FUNC[0], rel(f0, PUBLISHED_FUNC[0])))
    assert rel(f0, PUBLISHED_FUNC[0]) < 5e-4
    dt = run["prob"]["mp"]["dt_data"]
    for T in (1, 2, 4, 8, 12, 16, 20, 24, 32, 40, 48, 64, 96, 128):
        fT = reference_functional_fp32(run["r0"], dt, T)
        print("  float-accumulated functional, %3d threads: %.6f (rel to published %.2e)"
              % (T, fT, rel(fT, PUBLISHED_FUNC[0])))


def test_first_step_length(run):
    g0 = run["g0"].astype(np.float64)
    clean = 1.0 / np.linalg.norm(g0)
    G = cpu_add_true_accumulation(run["per_shot"], run["start"], run["prob"]["dx"], run["prob"]["dz"])
    quirk = 1.0 / np.linalg.norm(G)
    gsum = np.sum(run["per_shot"], axis=0)
    Gn = cpu_add_true_accumulation(run["per_shot_nolag"], run["start"], run["prob"]["dx"], run["prob"]["dz"])
    print("first step a0/|g0|: clean %.6f (shot sum %.6f), CPU add=true accumulation %.6f, published %.6f"
          % (clean, 1.0 / np.linalg.norm(gsum), quirk, PUBLISHED_TRIALS[0][2]))
    print("  without the imaging lag: clean %.6f, CPU add=true accumulation %.6f"
          % (1.0 / np.linalg.norm(np.sum(run["per_shot_nolag"], axis=0)), 1.0 / np.linalg.norm(Gn)))
    assert np.linalg.norm(gsum - g0) / np.linalg.norm(g0) < 1e-5
    assert min(rel(clean, PUBLISHED_TRIALS[0][2]), rel(quirk, PUBLISHED_TRIALS[0][2])) < 5e-4


def report(name, hist, info):
    print("%s: iter  functional  published  rel" % name)
    for i, f in enumerate(hist):
        pub = PUBLISHED_FUNC[i] if i < len(PUBLISHED_FUNC) else float("nan")
        print("  %4d  %10.6f  %10.6f  %.2e" % (i, f, pub, rel(f, pub)))
    print("  trials: iteration trial step f |g| | published step f |g|")
    for k, row in enumerate(info["trials"]):
        pub = PUBLISHED_TRIALS[k] if k < len(PUBLISHED_TRIALS) else None
        print("  %d %d %.6f %.6f %s | %s" % (row[0], row[1], row[2], row[3],
                                             "%.6f" % row[4] if row[4] == row[4] else "-", pub))
    print("  feval %d geval %d (published %d %d)" % (info["feval"], info["geval"], PUBLISHED_FEVAL, PUBLISHED_GEVAL))


def test_lbfgs_trace(run):
    """The product's gradient (the shot sum, no add=true re-transform): the same search
    pattern and evaluation counts as the published run."""
    hist, info = run["hist"], run["info"]
    report("clean gradient", hist, info)
    assert run["iters"] == 10
    assert np.all(np.diff(hist) < 0)
    assert (info["feval"], info["geval"]) == (PUBLISHED_FEVAL, PUBLISHED_GEVAL)
    assert [tuple(r[:2].astype(int)) for r in info["trials"]] == [t[:2] for t in PUBLISHED_TRIALS]


def test_lbfgs_trace_cpu_build(run):
    """With the CPU build's add=true accumulation emulated on the per-shot gradients, the
    published trace: every trial's step length, function value and gradient norm."""
    hist, info = run["cpu_hist"], run["cpu_info"]
    report("CPU-build accumulation", hist, info)
    assert run["cpu_iters"] == 10
    assert (info["feval"], info["geval"]) == (PUBLISHED_FEVAL, PUBLISHED_GEVAL)
    trials = info["trials"]
    assert len(trials) == len(PUBLISHED_TRIALS)
    for row, pub in zip(trials, PUBLISHED_TRIALS):
        assert tuple(row[:2].astype(int)) == pub[:2]
