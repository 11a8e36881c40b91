// Batched acoustic propagation engine: nl_we_op_a::propagate (src/cpp/we_op.cpp:1937-2135)
// for every shot of an acquisition at once, on one device.
//
// One pressure-like field per shot, [shot][nx][ldz] with ldz = nz rounded up to a quad;
// model (K, 1/rho) = (rho vp^2, 1/rho) as convert_model (we_op.cpp:1863-1881) produces
// it.  A time step is one launch over 32 x 32 tiles of every shot (k_astep), then the
// point sources (k_ainject).  The forward run keeps every sub-th level on the device (the
// reference's full wavefield) when that history fits the snapshot budget, and otherwise
// checkpoints of segments of about sqrt(2 nsnap) snapshots from which the adjoint run
// recomputes them, one segment at a time; the adjoint run images them (k_agrad) the way
// the reference's compute_gradients does, into one gradient slab per shot that are
// summed at the end.  The reference never ran this operator on a GPU.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_common.h"
#include "fwi2d_hip.h"
#include "kernels.h"
#include "trace_matrix.h"

namespace fwi2d {

extern thread_local std::string g_last_error;
void upload_constants(hipStream_t s);

namespace {

// the Dirichlet SAT parameter at free_surface_stiffness 1 (we_op.cpp:1962)
constexpr double kFreeSurfaceA = 0.2508560249;

struct Acoustic {
    fwi2d_grid g{};
    int device = 0;
    double alpha = kFreeSurfaceA;
    int ldz = 0, ntx = 0, ntz = 0;
    long long nxz = 0;
    hipStream_t st = nullptr;
    int ns = 0, nr = 0;
    bool model_ready = false;
    long long model_epoch = 0, snap_epoch = -1;
    DevMem d_K, d_ri, d_kdt2, d_div, d_tx, d_tz;
    DevMem d_lev, d_full, d_gslab, d_gout, d_src, d_rcv, d_adj;
    // snapshot segments: checkpoints, the segment buffer, recompute levels, the forward
    // run's source traces
    DevMem d_ckpt, d_seg, d_rlev, d_fsrc;
    long long snapshot_budget_mib = 131072;  // whole snapshot history up to this size (128 GiB)
    // 1: the adjoint step kernel images the snapshots; 0 (default): k_agrad before the step,
    // faster: fused, the step kernel holds 151 VGPRs and 42 KB of LDS, 3 wavefronts per SIMD,
    // and a 64-shot 500 x 1700 adjoint step with imaging takes 566 us against 211 + 300 us
    int image_fused = 0;
    bool whole = true;
    int seg_K = 0, nseg = 0, seg_loaded = -1;
    long long stat_recompute_steps = 0;
    InjList inj_src, inj_rcv;
    std::vector<PointEntry> ext;
    DevMem d_ext;

    Acoustic(const fwi2d_grid& grid, double stiffness, int dev) : g(grid), device(dev)
    {
        if (g.nx < 12 || g.nz < 12) throw std::runtime_error("the acoustic operator needs nx, nz >= 12");
        if (g.nt < 2 || !(g.dt > 0) || !(g.dx > 0) || !(g.dz > 0)) throw std::runtime_error("invalid grid");
        if (g.sub < 0) throw std::runtime_error("sub must be >= 0");
        const int bc[4] = {g.bc_top, g.bc_bottom, g.bc_left, g.bc_right};
        for (int k = 0; k < 4; k++)
            if (bc[k] < 1 || bc[k] > 3)
                throw std::runtime_error("acoustic boundaries: 1 (free surface), 2 (absorbing) or 3 (rigid)");
        // the reference raises the stiffness to at least 1 for this operator (we_op.cpp:16)
        alpha = kFreeSurfaceA / std::max(stiffness, 1.0);
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        upload_constants(st);
        ldz = (g.nz + 31) / 32 * 32;  // columns start on 128-byte lines
        nxz = (long long)g.nx * ldz;
        ntx = (g.nx + kAcTX - 1) / kAcTX;
        ntz = (ldz + kAcTZ - 1) / kAcTZ;
        std::vector<float> tz = taper_axis(g.nz, g.taper_top, g.taper_bottom, g.taper_strength);
        tz.resize(ldz, 0.f);
        std::vector<float> tx = taper_axis(g.nx, g.taper_left, g.taper_right, g.taper_strength);
        up(d_tz, tz.data(), tz.size() * sizeof(float));
        up(d_tx, tx.data(), tx.size() * sizeof(float));
        HIPCHECK(hipStreamSynchronize(st));
    }
    ~Acoustic()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }

    void up(DevMem& d, const void* h, size_t b)
    {
        d.reserve(std::max<size_t>(b, 16));
        if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, st));
    }

    // model (K, 1/rho) [2][nx][nz]; the leapfrog factor dt^2 K and the divisors of the
    // absorbing sides (asat_scale_boundaries, spatial_operators.cpp:779-860: a boundary
    // cell takes v dt / (2 h H0) from each absorbing side it lies on)
    void set_model(const float* m)
    {
        const int nx = g.nx, nz = g.nz;
        std::vector<float> K(nxz, 0.f), ri(nxz, 0.f), kd(nxz, 0.f), dv(nxz, 1.f);
        const bool ab[4] = {g.bc_top == 2, g.bc_bottom == 2, g.bc_left == 2, g.bc_right == 2};
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++) {
                const long long i = (long long)ix * ldz + iz;
                const float k = m[(long long)ix * nz + iz], r = m[(long long)nx * nz + (long long)ix * nz + iz];
                if (!(k > 0) || !(r > 0)) throw std::runtime_error("K and 1/rho must be positive");
                K[i] = k;
                ri[i] = r;
                kd[i] = (float)(g.dt * g.dt * (double)k);
                const double v = std::sqrt((double)k * r) * g.dt, h0 = kHWeights[0];
                double s = 0;
                if (ab[0] && iz == 0) s += v / (2 * g.dz * h0);
                if (ab[1] && iz == nz - 1) s += v / (2 * g.dz * h0);
                if (ab[2] && ix == 0) s += v / (2 * g.dx * h0);
                if (ab[3] && ix == nx - 1) s += v / (2 * g.dx * h0);
                dv[i] = (float)(1 + s);
            }
        up(d_K, K.data(), K.size() * sizeof(float));
        up(d_ri, ri.data(), ri.size() * sizeof(float));
        up(d_kdt2, kd.data(), kd.size() * sizeof(float));
        up(d_div, dv.data(), dv.size() * sizeof(float));
        HIPCHECK(hipStreamSynchronize(st));
        model_ready = true;
        model_epoch++;
    }

    // sources and receivers with delta_m3 footprints (injector.cpp:3-64): the sources
    // inject into the forward run, the receivers extract from it and inject the adjoint
    // sources into the adjoint run; source trace s and receiver trace r at offset s nt, r nt
    void set_shots(int nshots, const double* sxz, const int* nrec, const double* rxz, double ox, double oz)
    {
        if (nshots < 1) throw std::runtime_error("need at least one shot");
        const int nt = g.nt;
        std::vector<PointEntry> se, re, xe;
        int r = 0;
        for (int s = 0; s < nshots; s++) {
            if (nrec[s] < 0) throw std::runtime_error("negative receiver count");
            const Pivot q = point_pivot(g.nx, g.nz, g.dx, g.dz, sxz[2 * s], sxz[2 * s + 1], ox, oz);
            PointEntry e{};
            e.xi0 = q.px - 1;
            e.zi0 = q.pz - 1;
            e.trace = e.tz = (long long)s * nt;
            e.shot = s;
            inj_weights(q.ax, q.px, g.nx, g.dx, false, e.wx);
            inj_weights(q.az, q.pz, g.nz, g.dz, false, e.wz);
            se.push_back(e);
            for (int k = 0; k < nrec[s]; k++, r++) {
                const Pivot p = point_pivot(g.nx, g.nz, g.dx, g.dz, rxz[2 * r], rxz[2 * r + 1], ox, oz);
                PointEntry f{};
                f.xi0 = p.px - 1;
                f.zi0 = p.pz - 1;
                f.trace = f.tz = (long long)r * nt;
                f.shot = s;
                inj_weights(p.ax, p.px, g.nx, g.dx, false, f.wx);
                inj_weights(p.az, p.pz, g.nz, g.dz, false, f.wz);
                re.push_back(f);
                ext_weights(p.ax, f.wx);
                ext_weights(p.az, f.wz);
                xe.push_back(f);
            }
        }
        ns = nshots;
        nr = r;
        inj_src.build(se, st);
        inj_rcv.build(re, st);
        ext.swap(xe);
        up(d_ext, ext.data(), ext.size() * sizeof(PointEntry));
        HIPCHECK(hipStreamSynchronize(st));
        snap_epoch = -1;
    }

    void require_ready() const
    {
        if (!model_ready) throw std::runtime_error("no model: call fwi2d_acoustic_set_model first");
        if (ns < 1) throw std::runtime_error("no acquisition: call fwi2d_acoustic_set_shots first");
    }

    AcousticArgs base_args() const
    {
        AcousticArgs a{};
        a.nx = g.nx;
        a.nz = g.nz;
        a.ldz = ldz;
        a.nshots = ns;
        a.ntx = ntx;
        a.ntz = ntz;
        // shots per step workgroup: up to 8 while the launch keeps 2048 or more workgroups
        // (64 shots of 500 x 1700: 8 shots 209 us a step, 16 shots 214 us)
        a.spb = (int)std::max(1LL, std::min(8LL, (long long)ntx * ntz * ns / 2048));
        // the rectangle of tiles whose cells are all interior along both axes (no closure
        // rows, no SAT terms): columns 6..nx-7, rows 6..nz-7
        auto interior = [](int n, int t, int ntiles, int& i0, int& i1) {
            i0 = (6 + t - 1) / t;
            i1 = std::min(ntiles, (n - 6) / t);  // tile i ends at (i + 1) t <= n - 6
            if (i1 <= i0) i0 = i1 = 0;
        };
        interior(g.nx, kAcTX, ntx, a.itx0, a.itx1);
        interior(g.nz, kAcTZ, ntz, a.itz0, a.itz1);
        if (a.itx1 == a.itx0 || a.itz1 == a.itz0) a.itx0 = a.itx1 = a.itz0 = a.itz1 = 0;
        a.bc[0] = g.bc_top;
        a.bc[1] = g.bc_bottom;
        a.bc[2] = g.bc_left;
        a.bc[3] = g.bc_right;
        a.rdx2 = (float)(1.0 / (g.dx * g.dx));
        a.rdz2 = (float)(1.0 / (g.dz * g.dz));
        for (int k = 0; k < 4; k++) {
            a.dir_x[k] = (float)(kBoundaryDeriv[k] / (kHWeights[k] * g.dx * g.dx));
            a.dir_z[k] = (float)(kBoundaryDeriv[k] / (kHWeights[k] * g.dz * g.dz));
        }
        a.pen_x = (float)(1.0 / (alpha * kHWeights[0] * g.dx * g.dx));
        a.pen_z = (float)(1.0 / (alpha * kHWeights[0] * g.dz * g.dz));
        a.neu_x = (float)(1.0 / (kHWeights[0] * g.dx * g.dx));
        a.neu_z = (float)(1.0 / (kHWeights[0] * g.dz * g.dz));
        a.abs_x = (float)(1.0 / (2 * g.dt * kHWeights[0] * g.dx));
        a.abs_z = (float)(1.0 / (2 * g.dt * kHWeights[0] * g.dz));
        a.K = d_K.as<float>();
        a.ri = d_ri.as<float>();
        a.kdt2 = d_kdt2.as<float>();
        a.div = d_div.as<float>();
        a.taper_x = d_tx.as<float>();
        a.taper_z = d_tz.as<float>();
        a.shot_stride = nxz;
        a.rdx = (float)(1.0 / g.dx);
        a.rdz = (float)(1.0 / g.dz);
        return a;
    }

    // snapshot slots 1 + nt/sub, as the reference allocates its full wavefield
    // (we_op.cpp:923); a slot past the last stored level stays zero
    int nsnap() const { return g.sub > 0 ? 1 + g.nt / g.sub : 0; }
    float* level(int k) { return d_lev.as<float>() + (long long)k * ns * nxz; }
    float* snap(int idx) { return d_full.as<float>() + (long long)idx * ns * nxz; }

    // Segments of K snapshots (K about sqrt(2 nsnap)); segment j owns snapshots jK ..
    // (j+1)K - 1 and its buffer holds jK - 1 .. (j+1)K (K + 2 slots), every snapshot its
    // imaging reads.  Segment j >= 1 restarts from the levels (curr, prev) at step
    // (jK - 1) sub, stored by the forward run, which also keeps the last segment's
    // snapshots; the adjoint run images the segments last to first, recomputing each.
    void plan_segments()
    {
        const int n = nsnap();
        const double history = (double)n * ns * nxz * sizeof(float);
        whole = history <= (double)snapshot_budget_mib * 1048576.0;
        seg_K = std::max(1, std::min(n, (int)std::ceil(std::sqrt(2.0 * n))));
        nseg = (n + seg_K - 1) / seg_K;
    }
    int seg_base(int j) const { return std::max(j * seg_K - 1, 0); }
    int seg_last(int j) const { return std::min((j + 1) * seg_K, nsnap() - 1); }
    float* slot(int j, int idx) { return d_seg.as<float>() + (long long)(idx - seg_base(j)) * ns * nxz; }
    float* ckpt(int k) { return d_ckpt.as<float>() + (long long)k * ns * nxz; }
    // snapshot idx as the imaging of a snapshot of segment j reads it (segment j loaded:
    // its buffer holds the neighbours jK - 1 and (j+1)K too)
    float* snapshot(int j, int idx) { return whole ? snap(idx) : slot(j, idx); }

    // one step it on levels (p, c, n): the step kernel, then the point sources; with
    // img_idx >= 0 (adjoint run, the image_fused option) the step kernel also images
    // snapshot img_idx against c
    void step(float* p, float* c, float* n, int it, const InjList& inj, const float* tr, int img_idx = -1)
    {
        AcousticArgs a = img_idx >= 0 ? image_args(img_idx) : base_args();
        a.img = img_idx >= 0;
        a.taper_prev = it > 0;
        a.prev = p;
        a.curr = c;
        a.next = n;
        launch_astep(a, st);
        launch_ainject(inj.d_cells.as<InjCell>(), (int)inj.cells.size(), inj.d_ptr.as<int>(),
                       inj.d_ctx.as<long long>(), inj.d_cw.as<float>(), tr, it, n, nxz, ldz, a.kdt2, a.div, a.taper_x,
                       a.taper_z, false, st);
    }

    // the initial condition prev = 1/2 K dt^2 inject(trace sample 0) (we_op.cpp:1977-1979)
    // on zeroed levels
    void init_levels(float* lev, const InjList& inj, const float* tr)
    {
        const long long ss = (long long)ns * nxz;
        launch_fill(lev, 3 * ss, 0.f, st);
        AcousticArgs a = base_args();
        launch_ainject(inj.d_cells.as<InjCell>(), (int)inj.cells.size(), inj.d_ptr.as<int>(),
                       inj.d_ctx.as<long long>(), inj.d_cw.as<float>(), tr, 0, lev, nxz, ldz, a.kdt2, a.div, a.taper_x,
                       a.taper_z, true, st);
    }

    // recompute segment j's snapshots from its checkpoint into the segment buffer
    void load_segment(int j)
    {
        if (seg_loaded == j) return;
        const long long ss = (long long)ns * nxz;
        const int sub = g.sub, lastw = (g.nt - 1) / sub;  // last written snapshot
        const int b = seg_base(j), e = seg_last(j);
        launch_fill(slot(j, b), (long long)(e - b + 1) * ss, 0.f, st);  // slots past lastw stay zero
        d_rlev.reserve((size_t)3 * ss * sizeof(float));
        float* R = d_rlev.as<float>();
        int ip = 0, ic = 1, in = 2, it = 0;
        if (j == 0) {
            init_levels(R, inj_src, d_fsrc.as<float>());
        } else {
            HIPCHECK(hipMemcpyAsync(R + ic * ss, ckpt(2 * (j - 1)), ss * sizeof(float), hipMemcpyDeviceToDevice, st));
            HIPCHECK(hipMemcpyAsync(R + ip * ss, ckpt(2 * (j - 1) + 1), ss * sizeof(float), hipMemcpyDeviceToDevice, st));
            it = b * sub;
        }
        const int end = std::min(e, lastw) * sub;
        for (;; it++) {
            if (it % sub == 0 && it / sub >= b)
                HIPCHECK(hipMemcpyAsync(slot(j, it / sub), R + ic * ss, ss * sizeof(float), hipMemcpyDeviceToDevice, st));
            if (it >= end) break;
            step(R + ip * ss, R + ic * ss, R + in * ss, it, inj_src, d_fsrc.as<float>());
            stat_recompute_steps++;
            const int t = ip;
            ip = ic;
            ic = in;
            in = t;
        }
        seg_loaded = j;
    }

    // one propagation (we_op.cpp:1937-2135): sources inj with traces tr; the forward run
    // extracts into rcv and keeps snapshots (or checkpoints), the adjoint run images them
    // into the slabs
    void propagate(bool adjoint, const InjList& inj, const float* tr, float* rcv, bool keep)
    {
        const int nt = g.nt, sub = g.sub;
        const long long ss = (long long)ns * nxz;
        d_lev.reserve((size_t)3 * ss * sizeof(float));
        init_levels(level(0), inj, tr);
        int ip = 0, ic = 1, in = 2;
        const int jlast = nseg - 1;
        auto capture = [&](int it) {  // forward run: snapshots and checkpoints of level it
            if (!keep) return;
            if (whole) {
                if (it % sub == 0)
                    HIPCHECK(hipMemcpyAsync(snap(it / sub), level(ic), ss * sizeof(float), hipMemcpyDeviceToDevice, st));
                return;
            }
            if (it % sub == 0 && it / sub >= seg_base(jlast))
                HIPCHECK(hipMemcpyAsync(slot(jlast, it / sub), level(ic), ss * sizeof(float), hipMemcpyDeviceToDevice,
                                        st));
            if (it % sub == 0 && (it / sub + 1) % seg_K == 0 && it / sub + 1 < nseg * seg_K && it < nt - 1) {
                const int j = (it / sub + 1) / seg_K;  // level (jK - 1) sub starts segment j
                if (j >= 1 && j < nseg) {
                    HIPCHECK(hipMemcpyAsync(ckpt(2 * (j - 1)), level(ic), ss * sizeof(float), hipMemcpyDeviceToDevice,
                                            st));
                    HIPCHECK(hipMemcpyAsync(ckpt(2 * (j - 1) + 1), level(ip), ss * sizeof(float),
                                            hipMemcpyDeviceToDevice, st));
                }
            }
        };
        for (int it = 0; it < nt - 1; it++) {
            int img_idx = -1;
            if (!adjoint) {
                capture(it);
                if (rcv) launch_aextract(level(ic), nxz, ldz, d_ext.as<PointEntry>(), nr, rcv, it, st);
            } else if ((nt - 1 - it) % sub == 0 && it != 0) {
                if (image_fused) img_idx = (nt - 1 - it) / sub;
                else image(level(ic), (nt - 1 - it) / sub);
            }
            step(level(ip), level(ic), level(in), it, inj, tr, img_idx);
            const int t = ip;
            ip = ic;
            ic = in;
            in = t;
        }
        if (!adjoint) {
            capture(nt - 1);
            if (rcv) launch_aextract(level(ic), nxz, ldz, d_ext.as<PointEntry>(), nr, rcv, nt - 1, st);
        } else {
            image(level(ic), 0);
        }
    }

    // compute_gradients (we_op.cpp:1883-1935) of snapshot idx with the adjoint level adj
    void image(const float* adj, int idx)
    {
        AcousticArgs a = image_args(idx);
        a.curr = adj;
        launch_agrad(a, st);
    }

    // the imaging arguments of snapshot idx (its segment loaded first)
    AcousticArgs image_args(int idx)
    {
        AcousticArgs a = base_args();
        const int last = g.nt / g.sub;
        const double dtg = g.sub * g.dt, t = idx * dtg;
        double w = 1.0;
        if (g.gdamp_tmax - t > 0 && t - g.gdamp_tmin >= 0)
            w = std::cos(g.gdamp_power * M_PI * (t - g.gdamp_tmin) / (g.gdamp_tmax - g.gdamp_tmin));
        w *= w;
        a.mode = (idx > 0 && idx < last) ? 0 : (idx == 0 ? 1 : 2);
        a.wk = (float)(w / dtg);
        a.wr = (float)(idx == 0 ? w * 0.5 * dtg : w * dtg);
        const int j = whole ? 0 : idx / seg_K;
        if (!whole) load_segment(j);
        a.f0 = snapshot(j, idx);
        a.fm = a.mode != 1 ? snapshot(j, idx - 1) : nullptr;
        a.fp = a.mode != 2 ? snapshot(j, idx + 1) : nullptr;
        a.fshot_stride = nxz;
        a.grad = d_gslab.as<float>();
        a.gslab_stride = 2 * nxz;
        return a;
    }

    // the operator level's per-trace time maps on the device (see trace_matrix.h)
    DevMem d_tm_ptr, d_tm_col, d_tm_w, d_tdata;
    void upload_matrix(const TraceMatrix& m, int ncol_expected, int nrow_expected)
    {
        check_trace_matrix(m, ncol_expected, nrow_expected);
        up(d_tm_ptr, m.ptr.data(), m.ptr.size() * sizeof(int));
        up(d_tm_col, m.col.data(), m.col.size() * sizeof(int));
        up(d_tm_w, m.w.data(), m.w.size() * sizeof(double));
    }

    void forward_traces(const float* src, bool keep, const TraceMatrix& m, float* out)
    {
        require_ready();
        upload_matrix(m, g.nt, -1);
        d_rcv.reserve(std::max<size_t>((size_t)nr * g.nt * sizeof(float), 16));
        up(d_src, src, (size_t)ns * g.nt * sizeof(float));
        forward(d_src.as<float>(), d_rcv.as<float>(), keep, true);
        d_tdata.reserve(std::max<size_t>((size_t)nr * m.nrow * sizeof(float), 16));
        launch_trace_matvec(d_tm_ptr.as<int>(), d_tm_col.as<int>(), d_tm_w.as<double>(), m.nrow, m.ncol,
                            d_rcv.as<float>(), d_tdata.as<float>(), nr, st);
        if ((long long)nr * m.nrow > 0)
            HIPCHECK(hipMemcpyAsync(out, d_tdata.p, (size_t)nr * m.nrow * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
    }

    void adjoint_traces(const float* data, const TraceMatrix& m, float* grad)
    {
        require_ready();
        upload_matrix(m, -1, g.nt);
        up(d_tdata, data, std::max<size_t>((size_t)nr * m.ncol * sizeof(float), 4));
        d_adj.reserve(std::max<size_t>((size_t)nr * g.nt * sizeof(float), 16));
        launch_trace_matvec(d_tm_ptr.as<int>(), d_tm_col.as<int>(), d_tm_w.as<double>(), m.nrow, m.ncol,
                            d_tdata.as<float>(), d_adj.as<float>(), nr, st);
        adjoint(d_adj.as<float>(), grad, true, false);
    }

    void forward(const float* src, float* rcv, bool keep, bool on_device)
    {
        require_ready();
        if (keep && g.sub <= 0) throw std::runtime_error("keeping snapshots needs sub > 0");
        const int nt = g.nt;
        const float* dsrc = src;
        if (!on_device) {
            up(d_src, src, (size_t)ns * nt * sizeof(float));
            dsrc = d_src.as<float>();
        }
        float* drcv = rcv;
        if (!on_device) {
            d_rcv.reserve(std::max<size_t>((size_t)nr * nt * sizeof(float), 16));
            drcv = d_rcv.as<float>();
        }
        if (keep) {
            plan_segments();
            const long long ss = (long long)ns * nxz;
            if (whole) {
                d_full.reserve((size_t)nsnap() * ss * sizeof(float));
                for (int idx = (nt - 1) / g.sub + 1; idx < nsnap(); idx++) launch_fill(snap(idx), ss, 0.f, st);
            } else {
                d_ckpt.reserve((size_t)std::max(2 * (nseg - 1), 1) * ss * sizeof(float));
                d_seg.reserve((size_t)(seg_K + 2) * ss * sizeof(float));
                const int jl = nseg - 1;
                launch_fill(slot(jl, seg_base(jl)), (long long)(seg_last(jl) - seg_base(jl) + 1) * ss, 0.f, st);
                // the recomputation re-runs the sources: keep their traces
                d_fsrc.reserve((size_t)ns * nt * sizeof(float));
                HIPCHECK(hipMemcpyAsync(d_fsrc.p, dsrc, (size_t)ns * nt * sizeof(float), hipMemcpyDeviceToDevice, st));
                seg_loaded = jl;
            }
        }
        propagate(false, inj_src, dsrc, nr > 0 ? drcv : nullptr, keep);
        if (!on_device && nr > 0)
            HIPCHECK(hipMemcpyAsync(rcv, drcv, (size_t)nr * nt * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
        snap_epoch = keep ? model_epoch : -1;
    }

    void adjoint(const float* adj, float* grad, bool adj_dev, bool grad_dev)
    {
        require_ready();
        if (g.sub <= 0) throw std::runtime_error("the adjoint needs sub > 0");
        if (snap_epoch != model_epoch)
            throw std::runtime_error("no snapshots of a forward run on the current model: run fwi2d_acoustic_forward "
                                     "with keep_snapshots first");
        const int nt = g.nt;
        const float* dadj = adj;  // not read without receivers
        if (!adj_dev && nr > 0) {
            up(d_adj, adj, (size_t)nr * nt * sizeof(float));
            dadj = d_adj.as<float>();
        }
        const int groups = (ns + base_args().spb - 1) / base_args().spb;  // one slab per shot group
        d_gslab.reserve((size_t)groups * 2 * nxz * sizeof(float));
        launch_fill(d_gslab.as<float>(), (long long)groups * 2 * nxz, 0.f, st);
        propagate(true, inj_rcv, dadj, nullptr, false);
        d_gout.reserve((size_t)2 * nxz * sizeof(float));
        launch_sum_shots(d_gslab.as<float>(), 2 * nxz, groups, 2 * nxz, d_gout.as<float>(), 0, st);
        // [2][nx][ldz] -> [2][nx][nz]
        HIPCHECK(hipMemcpy2DAsync(grad, (size_t)g.nz * sizeof(float), d_gout.p, (size_t)ldz * sizeof(float),
                                  (size_t)g.nz * sizeof(float), (size_t)2 * g.nx,
                                  grad_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
    }
};

template <class F>
int guarded_ac(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return 0;
    } catch (const std::exception& ex) {
        g_last_error = ex.what();
        return 1;
    }
}

}  // namespace
}  // namespace fwi2d

struct fwi2d_acoustic {
    fwi2d::Acoustic* impl;
};

namespace fwi2d {

void acoustic_forward_traces(fwi2d_acoustic* e, const float* src, bool keep, const TraceMatrix& m, float* out)
{
    HIPCHECK(hipSetDevice(e->impl->device));
    e->impl->forward_traces(src, keep, m, out);
}

void acoustic_adjoint_traces(fwi2d_acoustic* e, const float* data, const TraceMatrix& m, float* grad)
{
    HIPCHECK(hipSetDevice(e->impl->device));
    e->impl->adjoint_traces(data, m, grad);
}

}  // namespace fwi2d

extern "C" {

int fwi2d_acoustic_create(const fwi2d_grid* grid, double free_surface_stiffness, int device, fwi2d_acoustic** out)
{
    return fwi2d::guarded_ac([&] {
        if (!grid || !out) throw std::runtime_error("null argument");
        *out = nullptr;
        fwi2d::Acoustic* a = new fwi2d::Acoustic(*grid, free_surface_stiffness, device);
        *out = new fwi2d_acoustic{a};
    });
}

int fwi2d_acoustic_set_model(fwi2d_acoustic* e, const float* model)
{
    return fwi2d::guarded_ac([&] {
        if (!e || !model) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->set_model(model);
    });
}

int fwi2d_acoustic_set_shots(fwi2d_acoustic* e, int nshots, const double* sxz, const int* nrec, const double* rxz,
                             double ox, double oz)
{
    return fwi2d::guarded_ac([&] {
        if (!e || !sxz || !nrec) throw std::runtime_error("null argument");
        int n = 0;
        for (int s = 0; s < nshots; s++) n += nrec[s];
        if (n > 0 && !rxz) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->set_shots(nshots, sxz, nrec, rxz, ox, oz);
    });
}

int fwi2d_acoustic_forward(fwi2d_acoustic* e, const float* src, float* rcv, int keep_snapshots, int on_device)
{
    return fwi2d::guarded_ac([&] {
        if (!e || !src || (!rcv && e->impl->nr > 0)) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->forward(src, rcv, keep_snapshots != 0, on_device != 0);
    });
}

int fwi2d_acoustic_adjoint(fwi2d_acoustic* e, const float* adj, float* grad, int on_device)
{
    return fwi2d::guarded_ac([&] {
        if (!e || !grad || (!adj && e->impl->nr > 0)) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->adjoint(adj, grad, on_device != 0, on_device != 0);
    });
}

int fwi2d_acoustic_set_option(fwi2d_acoustic* e, const char* key, long long value)
{
    return fwi2d::guarded_ac([&] {
        if (!e || !key) throw std::runtime_error("null argument");
        const std::string k = key;
        if (k == "snapshot_budget") {
            if (value < 0) throw std::runtime_error("snapshot_budget must be >= 0 (MiB)");
            e->impl->snapshot_budget_mib = value;
            e->impl->snap_epoch = -1;  // the next forward run plans the snapshots again
        } else if (k == "image_fused") {
            if (value != 0 && value != 1) throw std::runtime_error("image_fused must be 0 or 1");
            e->impl->image_fused = (int)value;
        } else {
            throw std::runtime_error("unknown acoustic engine option: " + k);
        }
    });
}

long long fwi2d_acoustic_get_stat(fwi2d_acoustic* e, const char* key)
{
    if (!e || !key) return -1;
    const std::string k = key;
    const fwi2d::Acoustic& a = *e->impl;
    if (k == "recompute_steps") return a.stat_recompute_steps;
    if (k == "whole_snapshots") return a.whole ? 1 : 0;
    if (k == "segment") return a.seg_K;
    if (k == "nseg") return a.nseg;
    if (k == "snapshot_budget") return a.snapshot_budget_mib;
    if (k == "image_fused") return a.image_fused;
    if (k == "shots_per_block") return a.base_args().spb;
    return -1;
}

void fwi2d_acoustic_destroy(fwi2d_acoustic* e)
{
    if (!e) return;
    (void)hipSetDevice(e->impl->device);
    delete e->impl;
    delete e;
}

}  // extern "C"
