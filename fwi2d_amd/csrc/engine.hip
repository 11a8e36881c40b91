// Batched elastic propagation engine: device state, shot tables, checkpointed adjoint.
//
// The engine replaces the reference's per-shot device hook propagate_gpu
// (src/cuda/cudaWeOp.cpp:10-876): instead of one host-driven shot with ~60 launches
// and ~20 synchronisations per step, every time step of all shots is one fused
// launch, and the adjoint recomputes forward snapshots from device checkpoints
// (segments of K snapshots) instead of streaming a host-stored full wavefield.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fwi2d_hip.h"
#include "kernels.h"
#include "sbp_tables.h"
#include "trace_matrix.h"
#include "engine_common.h"

namespace fwi2d {

thread_local std::string g_last_error;

// Block partials of the FP64 dot product (fwi2d_engine_dot_f64 and its allocation-free
// variant use the same count, so their sums are bit-identical)
constexpr int kDotBlocks = 1024;

struct Engine {
    fwi2d_grid g{};
    int device = 0;
    hipStream_t st = nullptr;
    int np = 3;
    int ldz = 0;          // device z pitch: nz rounded up to a multiple of 32 cells (128 bytes)
    long long nxz = 0;    // nx * ldz, the device size of one field
    int tile_x = 16, ntx = 0, ntz = 0;
    std::vector<int4> boxes;                        // {sx, ex, sz, ez} of every step box
    std::vector<int> list_core, list_bz, list_bz8, list_bx;  // box indices per box class
    DevMem d_boxes, d_list_core, d_list_bz, d_list_bz8, d_list_bx;
    int opt_tile_x = 16, opt_segment = 0, opt_lag = 1, opt_graph = 0, opt_timing = 0;
    long long opt_ckpt_budget_mb = 0;  // device memory for checkpoints and snapshot slots; 0: sqrt scheme
    // hipGraph replay of whole time loops: a loop is captured once per configuration
    // (every pointer and size it uses) and replayed while that configuration holds
    struct LoopGraph {
        std::vector<long long> key;
        hipGraphExec_t exec = nullptr;
        int seg_loaded_after = -1;
        long long steps = 0, recompute = 0, images = 0;
        void reset()
        {
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
            key.clear();
        }
    };
    LoopGraph graph_fwd, graph_adj;
    long long graph_captures = 0, graph_replays = 0;
    int opt_autotune = 0;  // pick tile_x and shots per workgroup by timing steps at the next forward run
    bool tuned = false;
    std::string tune_report;
    int opt_spb = 8;       // shots advanced one after the other by each core-box workgroup
    int opt_spb_edge = 1;  // the same for edge tiles (their workgroups are latency-bound)
    int opt_spb_edge_img = 0;  // edge tiles of adjoint steps with the imaging fused in (0: opt_spb_edge)
    int opt_band_rows = 8;     // rows of the top band (the bottom band: 8-11 more than band_rows - 8)
    int opt_spb_img = 0;   // shots per imaging workgroup (0: 4, or opt_spb when fused)
    // 1: imaging condition fused into the adjoint step kernels (by parts in time, sub = 1);
    // 0: the imaging kernel after each adjoint step; -1: fused on isotropic grids, where it
    // is faster (the VTI fused kernels exceed the register file)
    int opt_image_fused = -1;
    bool image_fused() const { return opt_image_fused < 0 ? !vti : opt_image_fused != 0; }
    int opt_img_buffers = 1;  // LDS window buffers of the imaging kernel (1 or 2)
    int opt_img_stream = 1;   // 1: imaging kernel on its own stream beside the next step
    int opt_img_sbp = 1;      // 1: density term by parts in time (sub = 1, separate imaging)
    int opt_edge_stream = 1;  // edge-band launches on their own stream (1: before the core launch, 2: after)
    hipStream_t st_edge = nullptr;       // edge-tile launches run beside the interior ones
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // imaging of adjoint step it runs on st_img beside step it+1; step it+2 (which
    // overwrites level it) and any recompute (which overwrites snapshot slots) wait for it
    hipStream_t st_img = nullptr;
    hipEvent_t ev_img_in = nullptr, ev_img_done[2] = {nullptr, nullptr};
    bool img_out[2] = {false, false};
    double last_fwd_loop_ms = 0.0;  // HIP-event time of the last forward step loop (timing option)
    double last_adj_ms = 0.0;       // HIP-event time of the last adjoint run, recomputation included
    int last_fwd_loop_steps = 0;
    bool vti = false;

    DevMem d_model, d_deriv, d_div_tb, d_div_lr, d_tx, d_tz;
    std::vector<float> h_model, h_deriv;  // h_deriv: dt^2/rho[, (1+2eps) mu, (1+2eps) lambda], [nd][nx][nz]
    long long model_epoch = 0, ckpt_epoch = -1;

    // acquisition
    int ns = 0, nr = 0;
    int src_nc = 2;        // source trace components: 2 (force x, z) or 3 (mxx, mzz, mxz)
    double gauge = 0.0;    // > 0: DAS receivers of this gauge length
    DevMem d_das_cs;       // DAS: cos, sin of each channel's dip (the operator level's strain)
    std::vector<PointEntry> src_ents, rcv_ents;  // flat, ordered by shot
    std::vector<int> src_shot_ptr, rcv_shot_ptr;
    DevMem d_src_ents, d_rcv_ents, d_src_shot_ptr, d_rcv_shot_ptr;
    InjList inj_src, inj_rcv;  // sources (forward) and receivers (adjoint) as injectors

    // state and checkpoints
    DevMem d_fwd, d_adj, d_ckpt, d_seg, d_gshot, d_gout, d_src, d_rcv, d_adjsrc, d_partial, d_scalar;
    DevMem d_ckpt16, d_ckpt_scale, d_amax;  // 16-bit checkpoint levels, their per-shot scales
    DevMem d_tm_ptr, d_tm_col, d_tm_w, d_tdata;  // operator-level trace matrix and data traces
    int opt_bf16 = 0;                       // checkpoint levels: 0 FP32, 1 BF16, 2 FP16 (max-abs scaled)
    int opt_stiff = 0;                      // VTI raw gradients w.r.t. c11, c13, c33, c55, rho
    int seg_K = 0, nseg = 0, nsnap = 0, seg_loaded = -1;
    const float* fwd_src = nullptr;  // source traces of the checkpointed forward run (device)
    long long stat_steps = 0, stat_recompute_steps = 0, stat_image_calls = 0;

    // ------------------------------------------------------------------
    explicit Engine(const fwi2d_grid& grid, int dev) : g(grid), device(dev)
    {
        if (g.nx < 12 || g.nz < 12) throw std::runtime_error("grid must have nx >= 12 and nz >= 12");
        if (g.nt < 2) throw std::runtime_error("nt must be >= 2");
        if (!(g.dx > 0 && g.dz > 0 && g.dt > 0)) throw std::runtime_error("dx, dz and dt must be positive");
        if (g.sub < 0) throw std::runtime_error("sub must be >= 0");
        if (g.sub > 1 && g.nt % g.sub == 0)
            throw std::runtime_error("sub > 1 with nt a multiple of sub leaves the last imaging snapshot undefined "
                                     "(the reference reads an unwritten snapshot); use nt = k*sub + 1");
        for (int b : {g.bc_top, g.bc_bottom, g.bc_left, g.bc_right})
            if (b != 1 && b != 2) throw std::runtime_error("boundary conditions must be 1 (free surface) or 2 (absorbing)");
        vti = g.vti != 0;
        np = vti ? 5 : 3;
        ldz = (g.nz + 31) & ~31;
        nxz = (long long)g.nx * ldz;
        if (nxz >= (1LL << 31)) throw std::runtime_error("grid too large: nx * nz must stay below 2^31 cells");
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHECK(hipStreamCreateWithFlags(&st_edge, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        HIPCHECK(hipStreamCreateWithFlags(&st_img, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&ev_img_in, hipEventDisableTiming));
        for (auto& e : ev_img_done) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        upload_constants(st);
        set_tile(opt_tile_x);
        build_tapers();
    }
    ~Engine()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        if (st_edge) {
            (void)hipStreamSynchronize(st_edge);
            (void)hipStreamDestroy(st_edge);
        }
        if (st_img) {
            (void)hipStreamSynchronize(st_img);
            (void)hipStreamDestroy(st_img);
        }
        if (ev_img_in) (void)hipEventDestroy(ev_img_in);
        for (auto& e : ev_img_done)
            if (e) (void)hipEventDestroy(e);
        graph_fwd.reset();
        graph_adj.reset();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
    }

    void set_tile(int tx)
    {
        if (tx != 16 && tx != 32) throw std::runtime_error("tile_x must be 16 or 32");
        tile_x = tx;
        ntx = tiles_x(g.nx, tile_x);
        ntz = tiles_z(ldz);
        build_boxes();
    }

    // Step decomposition: the core [6, nx-6) x [8, core_z_end) where every stencil is
    // interior, cut into tile_x x 64 boxes (the tile grid clipped to the core), and the
    // edge bands around it: 120 x 8 / 64 x 16 boxes along the top and bottom (corners included),
    // 8 x 128 boxes along the left and right sides.  Boxes partition the grid.  The
    // band_rows option moves the core's top and bottom by band_rows - 8 rows into the
    // grid (thicker bands in 64 x 16 boxes).
    void build_boxes()
    {
        boxes.clear();
        list_core.clear();
        list_bz.clear();
        list_bz8.clear();
        list_bx.clear();
        const int nx = g.nx, nz = g.nz;
        const int cx0 = kCoreX0, cx1 = nx - kCoreX0;
        const int cz0 = std::min(kCoreZ0 + (opt_band_rows - 8), nz);
        const int cz1 = std::min(std::max(core_z_end(nz) - (opt_band_rows - 8), cz0), nz);
        auto add = [&](std::vector<int>& list, int sx, int ex, int sz, int ez) {
            if (ex <= sx || ez <= sz) return;
            list.push_back((int)boxes.size());
            boxes.push_back(make_int4(sx, ex, sz, ez));
        };
        for (int tix = 0; tix < ntx; tix++) {
            int sx, ex;
            tile_range_x(tix, ntx, tile_x, nx, sx, ex);
            for (int tiz = 0; tiz < ntz; tiz++) {
                int sz, ez;
                tile_range_z(tiz, ntz, nz, ldz, sz, ez);
                add(list_core, std::max(sx, cx0), std::min(ex, cx1), std::max(sz, cz0), std::min(ez, cz1));
            }
        }
        // top and bottom bands: 8-row bands in 120 x 8 boxes, taller ones (the bottom band
        // holds 8 to 11 rows with the default band_rows) in 64 x 16 boxes; a band of more
        // than 16 rows in rows of boxes split evenly (8 to 16 rows each).  Only a box at the
        // grid's top or bottom takes the thin class, whose compact window serves no other.
        auto zband = [&](int sz0, int ez) {
            const int n = ez - sz0, nchunk = (n + 15) / 16;
            for (int c = 0; c < nchunk; c++) {
                const int sz = sz0 + n * c / nchunk, e = sz0 + n * (c + 1) / nchunk;
                const bool thin = e - sz <= 8 && (sz == 0 || e == nz);
                const int w = thin ? kBandZ8Width : 64;
                const int nb = (nx + w - 1) / w;
                for (int k = 0; k < nb; k++) {
                    int sx, ex;
                    tile_range_x(k, nb, w, nx, sx, ex);
                    add(thin ? list_bz8 : list_bz, sx, ex, sz, e);
                }
            }
        };
        zband(0, cz0);
        zband(cz1, nz);
        for (int sz = cz0; sz < cz1; sz += 128) {
            const int ez = std::min(sz + 128, cz1);
            add(list_bx, 0, cx0, sz, ez);
            add(list_bx, cx1, nx, sz, ez);
        }
        auto up = [&](DevMem& d, const void* h, size_t b) {
            d.reserve(std::max<size_t>(b, 16));
            if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, st));
        };
        up(d_boxes, boxes.data(), boxes.size() * sizeof(int4));
        up(d_list_core, list_core.data(), list_core.size() * sizeof(int));
        up(d_list_bz, list_bz.data(), list_bz.size() * sizeof(int));
        up(d_list_bz8, list_bz8.data(), list_bz8.size() * sizeof(int));
        up(d_list_bx, list_bx.data(), list_bx.size() * sizeof(int));
        HIPCHECK(hipStreamSynchronize(st));
    }
    int ntiles() const { return (int)boxes.size(); }

    // cos^2 taper factors per axis (taperz/taperx, spatial_operators.cpp:523-549)
    void build_tapers()
    {
        std::vector<float> tz = taper_axis(g.nz, g.taper_top, g.taper_bottom, g.taper_strength);
        tz.resize(ldz, 0.f);
        std::vector<float> tx = taper_axis(g.nx, g.taper_left, g.taper_right, g.taper_strength);
        d_tz.reserve(tz.size() * sizeof(float));
        d_tx.reserve(tx.size() * sizeof(float));
        HIPCHECK(hipMemcpyAsync(d_tz.p, tz.data(), tz.size() * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(d_tx.p, tx.data(), tx.size() * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    }

    // Divisors 1 + c dt/(2 h h0) of the absorbing boundaries, with the corner
    // accumulation of esat_scale_boundaries / vtisat_scale_boundaries
    // (spatial_operators.cpp:307-521) for the full grid.
    void build_divisors()
    {
        const int nx = g.nx, nz = g.nz;
        const long long nc = (long long)nx * nz;  // host model is [np][nx][nz]
        const float* la = h_model.data();
        const float* mu = la + nc;
        const float* rho = la + 2 * nc;
        const float* eps = vti ? la + 4 * nc : nullptr;
        std::vector<double> dvx(nc, 1.0), dvz(nc, 1.0);
        const double h0 = kH0, dt = g.dt, dx = g.dx, dz = g.dz;
        auto vpx2 = [&](long long i) { return (vti ? (1 + 2 * (double)eps[i]) : 1.0) * ((double)la[i] + 2 * (double)mu[i]); };
        auto scale = [&](long long i, double sx, double sz) { dvx[i] = 1 + sx; dvz[i] = 1 + sz; };
        const bool top = g.bc_top == 2, bottom = g.bc_bottom == 2, left = g.bc_left == 2, right = g.bc_right == 2;
        double sc1x = 0, sc1z = 0, sc2x = 0, sc2z = 0;
        if (top) {
            sc1x = std::sqrt((double)mu[0] / rho[0]) * dt / (2 * dz * h0);
            sc1z = sc1x * std::sqrt(((double)la[0] + 2 * (double)mu[0]) / mu[0]);
            long long tr = (long long)(nx - 1) * nz;
            sc2x = std::sqrt((double)mu[tr] / rho[tr]) * dt / (2 * dz * h0);
            sc2z = sc1x * std::sqrt(((double)la[tr] + 2 * (double)mu[tr]) / mu[tr]);  // reference uses sc1x here
            for (int ix = 1; ix < nx - 1; ix++) {
                long long i = (long long)ix * nz;
                double sx = std::sqrt((double)mu[i] / rho[i]) * dt / (2 * dz * h0);
                scale(i, sx, sx * std::sqrt(((double)la[i] + 2 * (double)mu[i]) / mu[i]));
            }
        }
        if (left) {
            sc1x += std::sqrt(vpx2(0) / rho[0]) * dt / (2 * dx * h0);
            sc1z += std::sqrt((double)mu[0] / rho[0]) * dt / (2 * dx * h0);
            scale(0, sc1x, sc1z);
            long long bl = nz - 1;
            sc1x = std::sqrt(vpx2(bl) / rho[bl]) * dt / (2 * dx * h0);
            sc1z = sc1x * std::sqrt((double)mu[bl] / vpx2(bl));
            for (int iz = 1; iz < nz - 1; iz++) {
                double sx = std::sqrt(vpx2(iz) / rho[iz]) * dt / (2 * dx * h0);
                scale(iz, sx, sx * std::sqrt((double)mu[iz] / vpx2(iz)));
            }
        } else {
            scale(0, sc1x, sc1z);
            sc1x = sc1z = 0;
        }
        if (bottom) {
            long long bl = nz - 1;
            sc1x += std::sqrt((double)mu[bl] / rho[bl]) * dt / (2 * dz * h0);
            sc1z += std::sqrt(((double)la[bl] + 2 * (double)mu[bl]) / rho[bl]) * dt / (2 * dz * h0);
            scale(bl, sc1x, sc1z);
            long long br = (long long)(nx - 1) * nz + nz - 1;
            sc1x = std::sqrt((double)mu[br] / rho[br]) * dt / (2 * dz * h0);
            sc1z = sc1x * std::sqrt(((double)la[br] + 2 * (double)mu[br]) / mu[br]);
            for (int ix = 1; ix < nx - 1; ix++) {
                long long i = (long long)ix * nz + nz - 1;
                double sx = std::sqrt((double)mu[i] / rho[i]) * dt / (2 * dz * h0);
                scale(i, sx, sx * std::sqrt(((double)la[i] + 2 * (double)mu[i]) / mu[i]));
            }
        } else {
            scale(nz - 1, sc1x, sc1z);
            sc1x = sc1z = 0;
        }
        if (right) {
            long long br = (long long)(nx - 1) * nz + nz - 1;
            sc1x += std::sqrt(vpx2(br) / rho[br]) * dt / (2 * dx * h0);
            sc1z += std::sqrt((double)mu[br] / rho[br]) * dt / (2 * dx * h0);
            scale(br, sc1x, sc1z);
            for (int iz = 1; iz < nz - 1; iz++) {
                long long i = (long long)(nx - 1) * nz + iz;
                double sx = std::sqrt(vpx2(i) / rho[i]) * dt / (2 * dx * h0);
                scale(i, sx, sx * std::sqrt((double)mu[i] / vpx2(i)));
            }
            long long tr = (long long)(nx - 1) * nz;
            sc2x += std::sqrt(vpx2(tr) / rho[tr]) * dt / (2 * dx * h0);
            sc2z += std::sqrt((double)mu[tr] / rho[tr]) * dt / (2 * dx * h0);
            scale(tr, sc2x, sc2z);
        } else {
            scale((long long)(nx - 1) * nz + nz - 1, sc1x, sc1z);
            scale((long long)(nx - 1) * nz, sc2x, sc2z);
        }
        // perimeter arrays: top row and bottom row (corners included), left and right columns
        std::vector<float> tb(4 * (size_t)nx), lr(4 * (size_t)nz);
        for (int ix = 0; ix < nx; ix++) {
            tb[ix] = (float)dvx[(long long)ix * nz];
            tb[nx + ix] = (float)dvz[(long long)ix * nz];
            tb[2 * nx + ix] = (float)dvx[(long long)ix * nz + nz - 1];
            tb[3 * nx + ix] = (float)dvz[(long long)ix * nz + nz - 1];
        }
        for (int iz = 0; iz < nz; iz++) {
            lr[iz] = (float)dvx[iz];
            lr[nz + iz] = (float)dvz[iz];
            lr[2 * nz + iz] = (float)dvx[(long long)(nx - 1) * nz + iz];
            lr[3 * nz + iz] = (float)dvz[(long long)(nx - 1) * nz + iz];
        }
        d_div_tb.reserve(tb.size() * sizeof(float));
        d_div_lr.reserve(lr.size() * sizeof(float));
        HIPCHECK(hipMemcpyAsync(d_div_tb.p, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(d_div_lr.p, lr.data(), lr.size() * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    }

    void set_model(const float* m)
    {
        const size_t n = (size_t)np * g.nx * g.nz;  // host model is [np][nx][nz]
        for (size_t i = 0; i < n; i++)
            if (!std::isfinite(m[i])) throw std::runtime_error("model contains a non-finite value");
        h_model.assign(m, m + n);
        const long long nc = (long long)g.nx * g.nz;
        const float* rho = m + 2 * nc;
        const float* mu = m + nc;
        for (long long i = 0; i < nc; i++)
            if (!(rho[i] > 0) || !(mu[i] > 0)) throw std::runtime_error("model must have rho > 0 and mu > 0");
        // device copy with z pitch ldz; padding cells hold lambda = mu = c13 = eps = 0, rho = 1
        d_model.reserve((size_t)np * nxz * sizeof(float));
        for (int c = 0; c < np; c++) launch_fill(d_model.as<float>() + c * nxz, nxz, c == 2 ? 1.f : 0.f, st);
        HIPCHECK(hipMemcpy2DAsync(d_model.p, (size_t)ldz * sizeof(float), m, (size_t)g.nz * sizeof(float),
                                  (size_t)g.nz * sizeof(float), (size_t)np * g.nx, hipMemcpyHostToDevice, st));
        // per-cell factors of the step, formed in double and rounded once
        // VTI imaging factors (we_op.cpp:1510-1519): delta recovered from c13, then
        // d c13 / d lambda and d c13 / d mu
        const int nd = vti ? 6 : 1;
        h_deriv.resize((size_t)nd * nc);
        for (long long i = 0; i < nc; i++) {
            h_deriv[i] = (float)(g.dt * g.dt / (double)rho[i]);
            if (vti) {
                const double l = m[i], u = mu[i], c = m[3 * nc + i];
                const double e = 1.0 + 2.0 * (double)m[4 * nc + i];
                h_deriv[nc + i] = (float)(e * u);
                h_deriv[2 * nc + i] = (float)(e * l);
                const double del = ((c + u) * (c + u) - (l + u) * (l + u)) / (2.0 * (l + u) * (l + 2.0 * u));
                const double v1 = std::sqrt(2.0 * (l + 2.0 * u) * (l + u) * del + (l + u) * (l + u));
                h_deriv[3 * nc + i] = (float)e;
                h_deriv[4 * nc + i] = (float)(((1.0 + 2.0 * del) * l + (1.0 + 3.0 * del) * u) / v1);
                h_deriv[5 * nc + i] = (float)(((1.0 + 3.0 * del) * l + (1.0 + 4.0 * del) * u) / v1 - 1.0);
            }
        }
        d_deriv.reserve((size_t)nd * nxz * sizeof(float));
        launch_fill(d_deriv.as<float>(), nd * nxz, 0.f, st);
        HIPCHECK(hipMemcpy2DAsync(d_deriv.p, (size_t)ldz * sizeof(float), h_deriv.data(), (size_t)g.nz * sizeof(float),
                                  (size_t)g.nz * sizeof(float), (size_t)nd * g.nx, hipMemcpyHostToDevice, st));
        build_divisors();
        model_epoch++;
    }

    // ------------------------------------------------------------------
    // acquisition: the 3x3 footprints of delta_m3 (injector.cpp:3-64), ddelta_m3
    // (injector.cpp:114-190) and dipole_m3 (injector.cpp:241-331)
    // ------------------------------------------------------------------
    Pivot pivot(double x, double z, double ox, double oz) const
    {
        return point_pivot(g.nx, g.nz, g.dx, g.dz, x, z, ox, oz);
    }
    PointEntry footprint(const Pivot& q, int shot, long long tx, long long tz) const
    {
        PointEntry e{};
        e.xi0 = q.px - 1;
        e.zi0 = q.pz - 1;
        e.trace = tx;
        e.tz = tz;
        e.shot = shot;
        return e;
    }

    std::vector<PointEntry> rcv_ext_ents;  // receiver footprints with extraction weights
    std::vector<int> rcv_ext_ptr;          // footprints of receiver r: rcv_ext_ptr[r] .. [r+1]
    DevMem d_rcv_ext_ents, d_rcv_ext_ptr;

    void set_shots(int nshots, const double* sxz, const int* nrec, const double* rxz, double ox, double oz, int mt = 0,
                   double gl = 0.0, const double* rdip = nullptr)
    {
        if (nshots < 1) throw std::runtime_error("need at least one shot");
        if (gl > 0 && !rdip) throw std::runtime_error("DAS receivers need their dips");
        int nrt = 0;
        for (int s = 0; s < nshots; s++) {
            if (nrec[s] < 0) throw std::runtime_error("negative receiver count");
            nrt += nrec[s];
        }
        const int nt = g.nt;
        std::vector<PointEntry> se, re, xe;
        std::vector<int> sp(nshots + 1, 0), rp(nshots + 1, 0), xp(1, 0);
        int r = 0;
        for (int s = 0; s < nshots; s++) {
            const Pivot q = pivot(sxz[2 * s], sxz[2 * s + 1], ox, oz);
            if (!mt) {  // point force: x and z traces of shot s
                PointEntry e = footprint(q, s, (long long)s * nt, (long long)(nshots + s) * nt);
                inj_weights(q.ax, q.px, g.nx, g.dx, false, e.wx);
                inj_weights(q.az, q.pz, g.nz, g.dz, false, e.wz);
                se.push_back(e);
            } else {
                // moment tensor [3][ns][nt] = (mxx, mzz, mxz): the force -d/dx (mxx, mxz) and
                // -d/dz (mxz, mzz) of the delta (we_op.cpp:640-654; injector.cpp:191-217)
                const long long txx = (long long)s * nt, tzz = (long long)(nshots + s) * nt,
                                txz = (long long)(2 * nshots + s) * nt;
                PointEntry a = footprint(q, s, txx, txz), b = footprint(q, s, txz, tzz);
                inj_weights(q.ax, q.px, g.nx, g.dx, true, a.wx);
                inj_weights(q.az, q.pz, g.nz, g.dz, false, a.wz);
                inj_weights(q.ax, q.px, g.nx, g.dx, false, b.wx);
                inj_weights(q.az, q.pz, g.nz, g.dz, true, b.wz);
                for (int i = 0; i < 3; i++) {
                    a.wx[i] = -a.wx[i];
                    b.wz[i] = -b.wz[i];
                }
                se.push_back(a);
                se.push_back(b);
            }
            sp[s + 1] = (int)se.size();
            for (int k = 0; k < nrec[s]; k++, r++) {
                const double x = rxz[2 * r], z = rxz[2 * r + 1];
                const long long tx = (long long)r * nt, tz = (long long)(nrt + r) * nt;
                if (gl <= 0) {
                    const Pivot p = pivot(x, z, ox, oz);
                    PointEntry e = footprint(p, s, tx, tz);
                    inj_weights(p.ax, p.px, g.nx, g.dx, false, e.wx);
                    inj_weights(p.az, p.pz, g.nz, g.dz, false, e.wz);
                    re.push_back(e);
                    ext_weights(p.ax, e.wx);
                    ext_weights(p.az, e.wz);
                    xe.push_back(e);
                } else {
                    // DAS channel: the gauge ends along the dip; injection -first + second end,
                    // extraction second - first end (injector.cpp:333-378)
                    const double c = std::cos(rdip[r]), sn = std::sin(rdip[r]);
                    const Pivot p1 = pivot(x - gl / 2 * c, z - gl / 2 * sn, ox, oz);
                    const Pivot p2 = pivot(x + gl / 2 * c, z + gl / 2 * sn, ox, oz);
                    PointEntry e1 = footprint(p1, s, tx, tz), e2 = footprint(p2, s, tx, tz);
                    inj_weights(p1.ax, p1.px, g.nx, g.dx, false, e1.wx);
                    inj_weights(p1.az, p1.pz, g.nz, g.dz, false, e1.wz);
                    inj_weights(p2.ax, p2.px, g.nx, g.dx, false, e2.wx);
                    inj_weights(p2.az, p2.pz, g.nz, g.dz, false, e2.wz);
                    for (int i = 0; i < 3; i++) e1.wx[i] = -e1.wx[i];
                    re.push_back(e1);
                    re.push_back(e2);
                    ext_weights(p2.ax, e2.wx);
                    ext_weights(p2.az, e2.wz);
                    ext_weights(p1.ax, e1.wx);
                    ext_weights(p1.az, e1.wz);
                    for (int i = 0; i < 3; i++) e1.wx[i] = -e1.wx[i];
                    xe.push_back(e2);
                    xe.push_back(e1);
                }
                xp.push_back((int)xe.size());
            }
            rp[s + 1] = (int)re.size();
        }
        ns = nshots;
        nr = nrt;
        src_nc = mt ? 3 : 2;
        gauge = gl > 0 ? gl : 0.0;
        if (gl > 0) {
            std::vector<double> cs(2 * (size_t)nrt);
            for (int k = 0; k < nrt; k++) {
                cs[2 * k] = std::cos(rdip[k]);
                cs[2 * k + 1] = std::sin(rdip[k]);
            }
            d_das_cs.reserve(std::max<size_t>(cs.size() * sizeof(double), 16));
            if (nrt) HIPCHECK(hipMemcpyAsync(d_das_cs.p, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
        src_ents.swap(se);
        rcv_ents.swap(re);
        rcv_ext_ents.swap(xe);
        rcv_ext_ptr.swap(xp);
        src_shot_ptr.swap(sp);
        rcv_shot_ptr.swap(rp);
        auto up = [&](DevMem& d, const void* h, size_t b) {
            d.reserve(std::max<size_t>(b, 16));
            if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, st));
        };
        up(d_src_ents, src_ents.data(), src_ents.size() * sizeof(PointEntry));
        up(d_rcv_ents, rcv_ents.data(), rcv_ents.size() * sizeof(PointEntry));
        up(d_rcv_ext_ents, rcv_ext_ents.data(), rcv_ext_ents.size() * sizeof(PointEntry));
        up(d_rcv_ext_ptr, rcv_ext_ptr.data(), rcv_ext_ptr.size() * sizeof(int));
        up(d_src_shot_ptr, src_shot_ptr.data(), src_shot_ptr.size() * sizeof(int));
        up(d_rcv_shot_ptr, rcv_shot_ptr.data(), rcv_shot_ptr.size() * sizeof(int));
        inj_src.build(src_ents, st);
        inj_rcv.build(rcv_ents, st);
        ckpt_epoch = -1;
        tuned = false;
        HIPCHECK(hipStreamSynchronize(st));
    }

    // ------------------------------------------------------------------
    // stepping
    // ------------------------------------------------------------------
    long long state_stride() const { return (long long)ns * 2 * nxz; }  // one time level of all shots

    StepArgs base_args() const
    {
        StepArgs a{};
        a.nx = g.nx;
        a.nz = g.nz;
        a.ldz = ldz;
        a.ntx = ntx;
        a.ntz = ntz;
        a.nshots = ns;
        a.nt = g.nt;
        a.rdx = (float)(1.0 / g.dx);
        a.rdz = (float)(1.0 / g.dz);
        a.rdx2 = (float)(1.0 / (g.dx * g.dx));
        a.rdz2 = (float)(1.0 / (g.dz * g.dz));
        a.c1x = (float)(kD1Interior[0] / g.dx);
        a.c2x = (float)(kD1Interior[1] / g.dx);
        a.c1z = (float)(kD1Interior[0] / g.dz);
        a.c2z = (float)(kD1Interior[1] / g.dz);
        a.dt = (float)g.dt;
        a.dt2 = (float)g.dt * (float)g.dt;
        a.sat_x = (float)(1.0 / (g.dx * kH0));
        a.sat_z = (float)(1.0 / (g.dz * kH0));
        a.bc_top = g.bc_top;
        a.bc_bottom = g.bc_bottom;
        a.bc_left = g.bc_left;
        a.bc_right = g.bc_right;
        const float* m = d_model.as<float>();
        a.lam = m;
        a.mu = m + nxz;
        a.rho = m + 2 * nxz;
        a.c13 = vti ? m + 3 * nxz : nullptr;
        a.eps = vti ? m + 4 * nxz : nullptr;
        const float* dv = d_deriv.as<float>();
        a.irho = dv;
        a.mux = vti ? dv + nxz : nullptr;
        a.lamx = vti ? dv + 2 * nxz : nullptr;
        a.spb = std::max(1, std::min(opt_spb, ns));
        a.ngroups = (ns + a.spb - 1) / a.spb;
        a.div_tb = d_div_tb.as<float>();
        a.div_lr = d_div_lr.as<float>();
        a.taper_x = d_tx.as<float>();
        a.taper_z = d_tz.as<float>();
        a.shot_stride = 2 * nxz;
        return a;
    }

    // prev/curr/next of one propagation (forward or adjoint)
    // prev/curr/next of one propagation (forward or adjoint), rotating over n = 3 levels,
    // or 4 when the imaging of a step still reads the step's previous level while the
    // next step runs
    struct Levels {
        float* b[4] = {nullptr, nullptr, nullptr, nullptr};
        int n = 3;
        int prev = 0, curr = 1, next = 2;
        void rotate()
        {
            prev = curr;
            curr = next;
            next = (next + 1) % n;
        }
        float* P() const { return b[prev]; }
        float* C() const { return b[curr]; }
        float* N() const { return b[next]; }
    };

    Levels levels_in(DevMem& m, int n = 3)
    {
        const long long ss = state_stride();
        m.reserve((size_t)n * ss * sizeof(float));
        Levels L;
        L.n = n;
        for (int i = 0; i < n; i++) L.b[i] = m.as<float>() + i * ss;
        return L;
    }

    // zero state, then prev = 1/2 dt^2/rho * inject(trace sample 0)
    void init_levels(Levels& L, bool adjoint, const float* traces)
    {
        const long long ss = state_stride();
        launch_fill(L.b[0], L.n * ss, 0.f, st);
        L.prev = 0;
        L.curr = 1;
        L.next = 2;
        if (adjoint)
            launch_init_prev(L.P(), 2 * nxz, nxz, ldz, ns, d_rcv_shot_ptr.as<int>(), d_rcv_ents.as<PointEntry>(), traces,
                             st);
        else
            launch_init_prev(L.P(), 2 * nxz, nxz, ldz, ns, d_src_shot_ptr.as<int>(), d_src_ents.as<PointEntry>(), traces,
                             st);
        launch_scale_prev(L.P(), d_model.as<float>() + 2 * nxz, (float)(0.5 * g.dt * g.dt), nxz, ns, st);
    }

    // img_idx >= 0: image forward snapshot img_idx in this adjoint step (its segment must
    // be loaded): the step's current level is the lagged adjoint level and its new level
    // the adjoint level of the reference's compute_gradients call
    // sbp: img_idx is snapshot j of the by-parts imaging (imaged after this step with the
    // step's previous, current and new levels, see sbp_image_args), or, with the imaging
    // fused into the step (image_fused), the density term of snapshot j and the stiffness
    // terms of snapshot j-1, whose derivatives meet those of the step's current level
    void step(Levels& L, int it, bool adjoint, const float* traces, float* rcv, int img_idx = -1, bool sbp = false)
    {
        StepArgs a = base_args();
        // imaging fused into the step kernels, or (default) a separate k_img pass after the
        // step and its point sources
        const bool fuse_img = img_idx >= 1 && image_fused() && sbp;
        StepArgs ia{};
        if (img_idx >= 0) {
            if (sbp) sbp_image_args(a, img_idx);
            else set_image_args(a, img_idx);
            if (fuse_img) {
                int mode;
                float wrho;
                image_weights(img_idx - 1, mode, a.wlm, wrho);
                a.fm = seg_slot(img_idx / seg_K, img_idx - 1);
                a.img_deriv = 1;
            }
            ia = a;
            // shots per imaging workgroup: the separate kernel is fastest with small
            // groups (more workgroups in flight; measured 4 < 16 < 64 at 64 shots)
            const int spbi = opt_spb_img > 0 ? opt_spb_img : (fuse_img ? opt_spb : 4);
            ia.spb = std::max(1, std::min(spbi, ns));
            ia.ngroups = (ns + ia.spb - 1) / ia.spb;
            if (fuse_img) a = ia;
        }
        a.it = it;
        a.taper_prev = it > 0;
        a.prev = L.P();
        a.curr = L.C();
        a.next = L.N();
        // the new level overwrites level it-2, which the imaging of step it-2 reads
        if (img_out[it & 1]) {
            HIPCHECK(hipStreamWaitEvent(st, ev_img_done[it & 1], 0));
            img_out[it & 1] = false;
        }
        // extraction of the current level (forward)
        if (rcv)
            launch_extract(L.C(), 2 * nxz, nxz, ldz, d_rcv_ext_ents.as<PointEntry>(), d_rcv_ext_ptr.as<int>(), nr, rcv,
                           it, st);
        // edge bands on their own stream, concurrent with the core boxes (disjoint cells),
        // or after them on the same stream
        StepArgs e = a;
        e.spb = std::max(1, std::min(fuse_img && opt_spb_edge_img > 0 ? opt_spb_edge_img : opt_spb_edge, ns));
        e.ngroups = (ns + e.spb - 1) / e.spb;
        const bool img = fuse_img;
        const int4* bx = d_boxes.as<int4>();
        // edge_stream 1: bands enqueued before the core launch, 2: after it (they then fill
        // the core launch's last, partial round of workgroups)
        hipStream_t se = opt_edge_stream ? st_edge : st;
        if (opt_edge_stream) HIPCHECK(hipEventRecord(ev_fork, st));
        auto core = [&] {
            launch_step(a, vti, tile_x == 32 ? BoxClass::core32 : BoxClass::core16, bx, d_list_core.as<int>(),
                        (int)list_core.size(), img, st);
        };
        if (opt_edge_stream == 2) core();
        if (opt_edge_stream) HIPCHECK(hipStreamWaitEvent(st_edge, ev_fork, 0));
        launch_step(e, vti, BoxClass::band_z8, bx, d_list_bz8.as<int>(), (int)list_bz8.size(), img, se);
        launch_step(e, vti, BoxClass::band_z, bx, d_list_bz.as<int>(), (int)list_bz.size(), img, se);
        launch_step(e, vti, BoxClass::band_x, bx, d_list_bx.as<int>(), (int)list_bx.size(), img, se);
        if (opt_edge_stream) HIPCHECK(hipEventRecord(ev_join, st_edge));
        if (opt_edge_stream != 2) core();
        if (opt_edge_stream) HIPCHECK(hipStreamWaitEvent(st, ev_join, 0));
        // point sources of this step into the new level
        const InjList& inj = adjoint ? inj_rcv : inj_src;
        InjImage im{};
        if (img) {
            im.on = 1;
            im.wn = a.wn;
            im.f0 = a.f0;
            im.fshot_stride = a.fshot_stride;
            im.grad = a.grad;
            im.gslab_stride = a.gslab_stride;
            im.rho_off = (long long)(a.stiff ? 4 : 2) * nxz;
        }
        launch_inject_cells(inj.d_cells.as<InjCell>(), (int)inj.cells.size(), inj.d_ptr.as<int>(),
                            inj.d_ctx.as<long long>(), inj.d_ctz.as<long long>(), inj.d_cw.as<float>(), traces, it,
                            L.N(), 2 * nxz, nxz, g.nx, g.nz, ldz,
                            d_deriv.as<float>(), d_tx.as<float>(), d_tz.as<float>(), d_div_tb.as<float>(),
                            d_div_lr.as<float>(), im, st);
        if (img_idx >= 0 && !fuse_img) {
            ia.aprev = L.P();
            ia.curr = L.C();
            ia.next = L.N();
            run_img(ia, it, sbp);
        }
        L.rotate();
        stat_steps++;
    }

    // ------------------------------------------------------------------
    // checkpoint plan: segment j holds snapshots idx in [max(jK-1,0), min((j+1)K, nsnap-1)]
    // and restarts from the state at level (jK-1)*sub (j >= 1) or from the start (j = 0)
    // ------------------------------------------------------------------
    void plan_segments()
    {
        nsnap = (g.nt - 1) / g.sub + 1;
        int K = opt_segment > 0 ? opt_segment : (int)std::ceil(std::sqrt(2.0 * nsnap));
        K = std::max(1, std::min(K, nsnap));
        if (opt_segment == 0 && opt_ckpt_budget_mb > 0) {
            // the longest segments whose checkpoints (2 levels per segment after the first)
            // and snapshot slots (K + 2 levels) fit the budget: the last segment is kept
            // whole by the forward run, so every snapshot beyond the first nsnap - K is
            // never recomputed
            const double level = (double)state_stride() * sizeof(float);
            const double ck_level = (double)state_stride() * (opt_bf16 ? 2 : 4);
            const double budget = (double)opt_ckpt_budget_mb * 1048576.0;
            for (int k = nsnap; k > K; k--) {
                const int nseg_k = (nsnap + k - 1) / k;
                if (2.0 * (nseg_k - 1) * ck_level + (k + 2.0) * level <= budget) {
                    // the same number of segments of balanced length, so that the last one
                    // (kept whole) is not a short remainder
                    K = (nsnap + nseg_k - 1) / nseg_k;
                    break;
                }
            }
        }
        seg_K = K;
        nseg = (nsnap + K - 1) / K;
    }
    int seg_base(int j) const { return std::max(j * seg_K - 1, 0); }
    int seg_last_idx(int j) const { return std::min((j + 1) * seg_K, nsnap - 1); }
    float* seg_slot(int j, int idx) { return d_seg.as<float>() + (long long)(idx - seg_base(j)) * state_stride(); }
    // checkpoints exist for segments j = 1 .. nseg-1
    float* ckpt_curr(int j) { return d_ckpt.as<float>() + (long long)(2 * (j - 1)) * state_stride(); }
    float* ckpt_prev(int j) { return d_ckpt.as<float>() + (long long)(2 * (j - 1) + 1) * state_stride(); }
    // checkpoint slot k (2(j-1) curr, 2(j-1)+1 prev): store a level / restore it
    void ckpt_store(int k, const float* u)
    {
        const long long ss = state_stride();
        if (opt_bf16)
            launch_pack16(opt_bf16 == 2, u, d_ckpt16.as<unsigned short>() + k * ss,
                          d_ckpt_scale.as<float>() + (long long)k * ns, d_amax.as<unsigned>(), 2 * nxz, ns, st);
        else
            launch_copy(d_ckpt.as<float>() + k * ss, u, ss, st);
    }
    void ckpt_load(int k, float* u)
    {
        const long long ss = state_stride();
        if (opt_bf16)
            launch_unpack16(opt_bf16 == 2, d_ckpt16.as<unsigned short>() + k * ss,
                            d_ckpt_scale.as<float>() + (long long)k * ns, u, 2 * nxz, ns, st);
        else
            launch_copy(u, d_ckpt.as<float>() + k * ss, ss, st);
    }

    // the trace matrix m on the device, its shape and indices checked against the traces
    void upload_matrix(const TraceMatrix& m, int ncol_expected, int nrow_expected)
    {
        check_trace_matrix(m, ncol_expected, nrow_expected);
        auto up = [&](DevMem& d, const void* h, size_t b) {
            d.reserve(std::max<size_t>(b, 16));
            if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, st));
        };
        up(d_tm_ptr, m.ptr.data(), m.ptr.size() * sizeof(int));
        up(d_tm_col, m.col.data(), m.col.size() * sizeof(int));
        up(d_tm_w, m.w.data(), m.w.size() * sizeof(double));
    }

    void forward_traces(const float* src, bool keep, const TraceMatrix& m, float* out)
    {
        require_ready();
        const long long ntr = 2LL * nr;
        upload_matrix(m, g.nt, -1);
        d_rcv.reserve((size_t)std::max<long long>(ntr * g.nt, 1) * sizeof(float));
        forward(src, d_rcv.as<float>(), keep, false, true);
        d_tdata.reserve((size_t)std::max<long long>(ntr * m.nrow, 1) * sizeof(float));
        launch_trace_matvec(d_tm_ptr.as<int>(), d_tm_col.as<int>(), d_tm_w.as<double>(), m.nrow, m.ncol,
                            d_rcv.as<float>(), d_tdata.as<float>(), ntr, st);
        // DAS: gauge-end differences -> strain; it commutes with the per-trace time map
        if (gauge > 0) launch_das_strain(d_tdata.as<float>(), d_das_cs.as<double>(), nr, m.nrow, gauge, false, st);
        if (ntr * m.nrow > 0)
            HIPCHECK(hipMemcpyAsync(out, d_tdata.p, (size_t)(ntr * m.nrow) * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
    }

    void adjoint_traces(const float* data, const TraceMatrix& m, float* grad)
    {
        require_ready();
        const long long ntr = 2LL * nr;
        upload_matrix(m, -1, g.nt);
        float* ddata = upload_traces(d_tdata, data, std::max<long long>(ntr * m.ncol, 1));
        // DAS: strain -> gauge-end differences (we_op.cpp:1130-1135)
        if (gauge > 0) launch_das_strain(ddata, d_das_cs.as<double>(), nr, m.ncol, gauge, true, st);
        d_adjsrc.reserve((size_t)std::max<long long>(ntr * g.nt, 1) * sizeof(float));
        launch_trace_matvec(d_tm_ptr.as<int>(), d_tm_col.as<int>(), d_tm_w.as<double>(), m.nrow, m.ncol, ddata,
                            d_adjsrc.as<float>(), ntr, st);
        adjoint(d_adjsrc.as<float>(), grad, true, false);
    }

    float* upload_traces(DevMem& d, const float* h, long long n)
    {
        d.reserve((size_t)n * sizeof(float));
        HIPCHECK(hipMemcpyAsync(d.p, h, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
        return d.as<float>();
    }

    void require_ready() const
    {
        if (ns == 0) throw std::runtime_error("no shots: call fwi2d_engine_set_shots first");
        if (model_epoch == 0) throw std::runtime_error("no model: call fwi2d_engine_set_model first");
    }

    // Every value a captured loop depends on besides the contents of device memory.
    std::vector<long long> loop_key(int kind, const void* p0, const void* p1, int flag) const
    {
        auto P = [](const void* p) { return (long long)(uintptr_t)p; };
        return {kind, P(p0), P(p1), flag, g.nt, ns, nr, tile_x, opt_spb, opt_spb_edge, opt_spb_edge_img, opt_lag, seg_K, nseg,
                opt_bf16, opt_stiff, opt_edge_stream, opt_spb_img, opt_image_fused, opt_img_buffers, opt_img_stream, opt_img_sbp, P(d_ckpt16.p), P(d_ckpt_scale.p), P(d_amax.p),
                P(d_fwd.p), P(d_adj.p), P(d_ckpt.p), P(d_seg.p), P(d_gshot.p), P(d_model.p), P(d_deriv.p),
                P(d_boxes.p), P(d_list_core.p), P(d_list_bz.p), P(d_list_bz8.p), P(d_list_bx.p), P(inj_src.d_cells.p), P(inj_src.d_ptr.p),
                P(inj_src.d_ctx.p), P(inj_src.d_ctz.p), P(inj_src.d_cw.p), P(inj_rcv.d_cells.p), P(inj_rcv.d_ptr.p),
                P(inj_rcv.d_ctx.p), P(inj_rcv.d_ctz.p), P(d_rcv_ext_ptr.p), src_nc,
                P(inj_rcv.d_cw.p), (long long)inj_src.cells.size(), (long long)inj_rcv.cells.size(), P(d_src_ents.p),
                P(d_rcv_ents.p), P(d_rcv_ext_ents.p), P(d_src_shot_ptr.p), P(d_rcv_shot_ptr.p), P(d_tx.p), P(d_tz.p),
                P(d_div_tb.p), P(d_div_lr.p), (long long)list_core.size(), (long long)list_bz.size(), (long long)list_bz8.size(),
                (long long)list_bx.size(), (long long)boxes.size()};
    }

    // Runs body (which only enqueues device work on st / st_edge) directly, or, with the
    // graph option, replays its capture when the key matches and captures it otherwise.
    template <class F>
    void run_loop(LoopGraph& G, const std::vector<long long>& key, F body)
    {
        if (!opt_graph) {
            body();
            return;
        }
        if (G.exec && G.key == key) {
            HIPCHECK(hipGraphLaunch(G.exec, st));
            seg_loaded = G.seg_loaded_after;
            stat_steps += G.steps;
            stat_recompute_steps += G.recompute;
            stat_image_calls += G.images;
            graph_replays++;
            return;
        }
        G.reset();
        const long long s0 = stat_steps, r0 = stat_recompute_steps, i0 = stat_image_calls;
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(st, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK(ie);
        G.exec = exec;
        G.key = key;
        G.seg_loaded_after = seg_loaded;
        G.steps = stat_steps - s0;
        G.recompute = stat_recompute_steps - r0;
        G.images = stat_image_calls - i0;
        graph_captures++;
        HIPCHECK(hipGraphLaunch(G.exec, st));
    }

    // Step-shape autotuner: times a few forward steps of the current shots for every
    // (tile_x, shots per core workgroup, shots per edge workgroup) and keeps the fastest.
    // Forward results do not depend on the choice (bit for bit); gradients only through
    // the order in which a workgroup sums its shots.
    void autotune(const float* dsrc)
    {
        const int warm = 2, timed = 4;
        const int nsteps = std::min(warm + timed, g.nt - 1);
        if (nsteps < 2) return;
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        double best = 1e300;
        int btx = opt_tile_x, bspb = opt_spb, bspe = opt_spb_edge;
        tune_report.clear();
        const long long steps0 = stat_steps;
        int bes = opt_edge_stream;
        double t_es2[2][5][3];  // the edge_stream = 2 time of each (tile, spb, spb_edge)
        for (auto& a : t_es2)
            for (auto& b : a)
                for (double& c : b) c = 1e300;
        for (int tx : {16, 32})
            for (int spb : {1, 2, 4, 8, 16}) {
                if (spb > 1 && spb > ns) continue;
                for (int spe : {1, 2, 4})
                    for (int es : {1, 2, 0}) {
                        if (spe > 1 && spe > ns) continue;
                        set_tile(tx);
                        opt_spb = spb;
                        opt_spb_edge = spe;
                        opt_edge_stream = es;
                        Levels L = levels_in(d_fwd);
                        init_levels(L, false, dsrc);
                        for (int it = 0; it < nsteps; it++) {
                            if (it == nsteps - timed) HIPCHECK(hipEventRecord(e0, st));
                            step(L, it, false, dsrc, nullptr);
                        }
                        HIPCHECK(hipEventRecord(e1, st));
                        HIPCHECK(hipEventSynchronize(e1));
                        float ms = 0.f;
                        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
                        const double us = 1e3 * ms / std::min(timed, nsteps);
                        char buf[112];
                        snprintf(buf, sizeof(buf), "%stile_x=%d spb=%d spb_edge=%d edge_stream=%d: %.1f us",
                                 tune_report.empty() ? "" : "; ", tx, spb, spe, es, us);
                        tune_report += buf;
                        if (es == 2) {
                            const int ispb = spb == 1 ? 0 : spb == 2 ? 1 : spb == 4 ? 2 : spb == 8 ? 3 : 4;
                            t_es2[tx == 32][ispb][spe == 1 ? 0 : spe == 2 ? 1 : 2] = us;
                        }
                        if (us < best) {
                            best = us;
                            btx = tx;
                            bspb = spb;
                            bspe = spe;
                            bes = es;
                        }
                    }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // the forward step cannot tell the edge-stream orders apart (within 1%), but the
        // graph-replayed gradient can: the core launch issued before the edge bands
        // (edge_stream 2) took 6.45 s per gradient of 64 shots on 500 x 1700 (nt 4000)
        // against 6.66 s with the bands first, so that order wins forward-step ties
        {
            const int ispb = bspb == 1 ? 0 : bspb == 2 ? 1 : bspb == 4 ? 2 : bspb == 8 ? 3 : 4;
            if (bes != 2 && t_es2[btx == 32][ispb][bspe == 1 ? 0 : bspe == 2 ? 1 : 2] <= 1.015 * best) bes = 2;
        }
        set_tile(btx);
        opt_tile_x = btx;
        opt_spb = bspb;
        opt_spb_edge = bspe;
        opt_edge_stream = bes;
        stat_steps = steps0;
        tuned = true;
    }

    void forward(const float* src, float* rcv, bool keep, bool src_dev, bool rcv_dev)
    {
        require_ready();
        if (keep && g.sub <= 0) throw std::runtime_error("checkpoints need sub > 0");
        ckpt_epoch = -1;  // any forward run replaces the recomputation sources
        const long long ntr_s = (long long)src_nc * ns * g.nt, ntr_r = (long long)2 * nr * g.nt;
        const float* dsrc = src_dev ? src : upload_traces(d_src, src, ntr_s);
        if (opt_autotune && !tuned) autotune(dsrc);
        float* drcv = rcv;
        if (!rcv_dev) {
            d_rcv.reserve((size_t)std::max<long long>(ntr_r, 1) * sizeof(float));
            drcv = d_rcv.as<float>();
        }
        if (nr > 0) launch_fill(drcv, ntr_r, 0.f, st);
        Levels L = levels_in(d_fwd);
        init_levels(L, false, dsrc);
        int jlast = -1;
        if (keep) {
            plan_segments();
            const size_t nslot = (size_t)std::max(2 * (nseg - 1), 1);
            if (opt_bf16) {
                d_ckpt16.reserve(nslot * state_stride() * sizeof(unsigned short));
                d_ckpt_scale.reserve(nslot * ns * sizeof(float));
                d_amax.reserve((size_t)ns * sizeof(unsigned));
            } else {
                d_ckpt.reserve(nslot * state_stride() * sizeof(float));
            }
            d_seg.reserve((size_t)(seg_K + 2) * state_stride() * sizeof(float));
            jlast = nseg - 1;
        }
        const int nt = g.nt;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (opt_timing) {
            HIPCHECK(hipEventCreate(&ev0));
            HIPCHECK(hipEventCreate(&ev1));
            HIPCHECK(hipEventRecord(ev0, st));
        }
        const Levels L0 = L;
        run_loop(graph_fwd, loop_key(0, dsrc, drcv, keep ? jlast : -1), [&] {
            Levels Lc = L0;
            for (int it = 0; it < nt - 1; it++) {
                if (keep) capture(Lc, it, jlast);
                step(Lc, it, false, dsrc, nr > 0 ? drcv : nullptr);
            }
        });
        for (int it = 0; it < nt - 1; it++) L.rotate();  // the loop's final level rotation
        if (opt_timing) {
            HIPCHECK(hipEventRecord(ev1, st));
            HIPCHECK(hipEventSynchronize(ev1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
            last_fwd_loop_ms = ms;
            last_fwd_loop_steps = nt - 1;
            (void)hipEventDestroy(ev0);
            (void)hipEventDestroy(ev1);
        }
        if (keep) capture(L, nt - 1, jlast);
        if (nr > 0)
            launch_extract(L.C(), 2 * nxz, nxz, ldz, d_rcv_ext_ents.as<PointEntry>(), d_rcv_ext_ptr.as<int>(), nr, drcv,
                           nt - 1, st);
        if (!rcv_dev && nr > 0)
            HIPCHECK(hipMemcpyAsync(rcv, drcv, (size_t)ntr_r * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
        if (keep) {
            ckpt_epoch = model_epoch;
            seg_loaded = jlast;
            fwd_src = dsrc;
        }
    }

    // during a forward run: store checkpoints and the last segment's snapshots
    void capture(const Levels& L, int level, int jlast)
    {
        const long long ss = state_stride();
        for (int j = 1; j < nseg; j++)
            if (level == (j * seg_K - 1) * g.sub && level < g.nt - 1) {
                ckpt_store(2 * (j - 1), L.C());
                ckpt_store(2 * (j - 1) + 1, L.P());
            }
        if (level % g.sub == 0) {
            int idx = level / g.sub;
            if (idx >= seg_base(jlast) && idx <= seg_last_idx(jlast)) launch_copy(seg_slot(jlast, idx), L.C(), ss, st);
        }
    }

    // recompute the snapshots of segment j from its checkpoint
    // imaging launch of adjoint step it, on st_img beside the next step (or on st)
    void run_img(const StepArgs& ia, int it, bool sbp)
    {
        if (opt_img_stream) {
            HIPCHECK(hipEventRecord(ev_img_in, st));
            HIPCHECK(hipStreamWaitEvent(st_img, ev_img_in, 0));
            launch_img(ia, vti, tile_x, opt_img_buffers, sbp, st_img);
            HIPCHECK(hipEventRecord(ev_img_done[it & 1], st_img));
            img_out[it & 1] = true;
        } else {
            launch_img(ia, vti, tile_x, opt_img_buffers, sbp, st);
        }
    }

    // The density term by parts in time (sub = 1).  The reference's density gradient is
    // sum over imaged snapshots idx = 0 .. nt-2 of wrho(idx) <a_idx, D2 F_idx>, with a_idx
    // the adjoint level nt-1-idx and D2 F_idx = F_idx+1 - 2 F_idx + F_idx-1 (F_1 - F_0 at
    // idx = 0).  Collected per snapshot j it is <F_j, c_j-1 - 2 c_j + c_j+1>, c = wrho a,
    // where only imaged idx contribute (c_j-1 for j >= 1, -2 c_j for 1 <= j <= nt-2 and
    // -c_0 for j = 0, c_j+1 for j <= nt-3).  Snapshot j = nt-1-it is imaged after adjoint
    // step it: a_j-1, a_j, a_j+1 are its new, current and previous levels, and its
    // stiffness terms pair the previous level's derivatives with F_j's (the reference's
    // lag), for j <= nt-2.  Snapshot 0 is imaged after the last step.
    void sbp_image_args(StepArgs& a, int j)
    {
        const int sj = j / seg_K;
        if (seg_loaded != sj) throw std::logic_error("imaging segment not loaded");
        const int last = g.nt - 2;  // last imaged snapshot
        auto wr = [&](int i) {
            int mode;
            float wlm, wrho;
            image_weights(i, mode, wlm, wrho);
            return wrho;
        };
        int mode;
        float wrho;
        image_weights(j, mode, a.wlm, wrho);
        a.img_mode = mode;
        a.wrho = wrho;
        a.wn = (j >= 1 && j - 1 <= last) ? wr(j - 1) : 0.f;
        a.wc = (j >= 1 && j <= last) ? -2.f * wr(j) : (j == 0 ? -wr(0) : 0.f);
        a.wp = (j + 1 <= last) ? wr(j + 1) : 0.f;
        a.img_deriv = j <= last ? 1 : 0;
        a.f0 = seg_slot(sj, j);
        a.fm = a.fp = nullptr;
        a.fshot_stride = 2 * nxz;
        a.grad = d_gshot.as<float>();
        a.gslab_stride = (long long)np * nxz;
        const float* dv = d_deriv.as<float>();
        a.ge = vti ? dv + 3 * nxz : nullptr;
        a.gv2 = vti ? dv + 4 * nxz : nullptr;
        a.gv3 = vti ? dv + 5 * nxz : nullptr;
        a.stiff = vti && opt_stiff;
    }

    // the main stream waits for every imaging launch still running on st_img
    void img_join()
    {
        for (int p = 0; p < 2; p++)
            if (img_out[p]) {
                HIPCHECK(hipStreamWaitEvent(st, ev_img_done[p], 0));
                img_out[p] = false;
            }
    }

    void load_segment(int j, const float* dsrc)
    {
        if (seg_loaded == j) return;
        img_join();  // the recompute overwrites snapshot slots the imaging reads
        const long long ss = state_stride();
        Levels F = levels_in(d_fwd);
        if (g.sub == 1) {
            // every level is a snapshot: step straight from slot to slot of the segment
            // buffer (the level before the first slot lives in a scratch level)
            const int b = seg_base(j), last = seg_last_idx(j);
            float* p0 = F.b[0];
            if (j == 0) {
                init_levels(F, false, dsrc);  // F.P() = b[0]: initial condition, F.C(): zero level 0
                launch_copy(seg_slot(j, 0), F.C(), ss, st);
            } else {
                ckpt_load(2 * (j - 1), seg_slot(j, b));
                ckpt_load(2 * (j - 1) + 1, p0);
            }
            for (int idx = b; idx < last; idx++) {
                Levels T;
                T.b[0] = idx == b ? p0 : seg_slot(j, idx - 1);
                T.b[1] = seg_slot(j, idx);
                T.b[2] = seg_slot(j, idx + 1);
                step(T, idx, false, dsrc, nullptr);
                stat_recompute_steps++;
            }
            seg_loaded = j;
            return;
        }
        int start;
        if (j == 0) {
            init_levels(F, false, dsrc);
            start = 0;
        } else {
            F.prev = 0;
            F.curr = 1;
            F.next = 2;
            ckpt_load(2 * (j - 1), F.C());
            ckpt_load(2 * (j - 1) + 1, F.P());
            start = seg_base(j) * g.sub;
        }
        const int end = seg_last_idx(j) * g.sub;
        for (int it = start;; it++) {
            if (it % g.sub == 0) {
                int idx = it / g.sub;
                if (idx >= seg_base(j) && idx <= seg_last_idx(j)) launch_copy(seg_slot(j, idx), F.C(), ss, st);
            }
            if (it >= end) break;
            step(F, it, false, dsrc, nullptr);
            stat_recompute_steps++;
        }
        seg_loaded = j;
    }

    // imaging weights of snapshot idx (we_op.cpp:527-561): gdamp window, endpoint halves
    void image_weights(int idx, int& mode, float& wlm, float& wrho) const
    {
        mode = (idx > 0 && idx < g.nt / g.sub) ? 0 : (idx == 0 ? 1 : 2);
        const double dtg = g.sub * g.dt;
        const double t = idx * dtg;
        double w = 1.0;
        if (g.gdamp_tmax - t > 0 && t - g.gdamp_tmin >= 0)
            w = std::cos(g.gdamp_power * M_PI * (t - g.gdamp_tmin) / (g.gdamp_tmax - g.gdamp_tmin));
        w *= w;
        wlm = (float)(idx == 0 ? w * 0.5 * dtg : w * dtg);
        wrho = (float)(w * 1.0 / dtg);
    }

    void set_image_args(StepArgs& a, int idx)
    {
        const int j = idx / seg_K;
        if (seg_loaded != j) throw std::logic_error("imaging segment not loaded");
        image_weights(idx, a.img_mode, a.wlm, a.wrho);
        a.f0 = seg_slot(j, idx);
        a.fm = (a.img_mode != 1) ? seg_slot(j, idx - 1) : nullptr;
        a.fp = (a.img_mode != 2) ? seg_slot(j, idx + 1) : nullptr;
        a.fshot_stride = 2 * nxz;
        a.grad = d_gshot.as<float>();
        a.gslab_stride = (long long)np * nxz;
        const float* dv = d_deriv.as<float>();
        a.ge = vti ? dv + 3 * nxz : nullptr;
        a.gv2 = vti ? dv + 4 * nxz : nullptr;
        a.gv3 = vti ? dv + 5 * nxz : nullptr;
        a.stiff = vti && opt_stiff;
    }

    void image(const Levels& A, int idx, const float* dsrc)
    {
        const int j = idx / seg_K;
        load_segment(j, dsrc);
        ImageArgs a{};
        a.nx = g.nx;
        a.nz = g.nz;
        a.ldz = ldz;
        a.ntx = ntx;
        a.ntz = ntz;
        a.nshots = ns;
        a.rdx = (float)(1.0 / g.dx);
        a.rdz = (float)(1.0 / g.dz);
        image_weights(idx, a.mode, a.wlm, a.wrho);
        const float* m = d_model.as<float>();
        a.lam = m;
        a.mu = m + nxz;
        a.c13 = vti ? m + 3 * nxz : nullptr;
        a.eps = vti ? m + 4 * nxz : nullptr;
        a.adj = A.C();
        a.adj_prev = A.P();
        a.shot_stride = 2 * nxz;
        a.f0 = seg_slot(j, idx);
        a.fm = (a.mode != 1) ? seg_slot(j, idx - 1) : nullptr;
        a.fp = (a.mode != 2) ? seg_slot(j, idx + 1) : nullptr;
        a.fshot_stride = 2 * nxz;
        a.grad = d_gshot.as<float>();
        a.ncomp = np;
        a.gshot_stride = (long long)np * nxz;
        a.lag = opt_lag;
        launch_image(a, vti, tile_x, st);
        stat_image_calls++;
    }

    void adjoint(const float* adj, float* grad, bool adj_dev, bool grad_dev)
    {
        require_ready();
        if (g.sub <= 0) throw std::runtime_error("the adjoint needs sub > 0");
        if (ckpt_epoch != model_epoch)
            throw std::runtime_error("no checkpoints of a forward run on the current model: run fwi2d_engine_forward "
                                     "with keep_checkpoints first");
        const long long ntr_r = (long long)2 * nr * g.nt;
        const float* dadj = adj_dev ? adj : upload_traces(d_adjsrc, adj, std::max<long long>(ntr_r, 1));
        const float* dsrc = fwd_src;
        // with the reference's lag the imaging is fused into the adjoint steps (each
        // workgroup adds its shots' sum to the slab of its first shot); without it (a test
        // option) the separate imaging kernel fills one slab per shot
        const bool fused = opt_lag == 1;
        if (opt_stiff && (!vti || !fused))
            throw std::runtime_error("stiffness gradients need a VTI grid and the imaging lag");
        const int nslab = ns;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (opt_timing) {
            HIPCHECK(hipEventCreate(&ev0));
            HIPCHECK(hipEventCreate(&ev1));
            HIPCHECK(hipEventRecord(ev0, st));
        }
        d_gshot.reserve((size_t)nslab * np * nxz * sizeof(float));
        launch_fill(d_gshot.as<float>(), (long long)nslab * np * nxz, 0.f, st);
        const int nt = g.nt, sub = g.sub;
        // the density term by parts needs each step's previous level while the next step
        // runs: four adjoint levels
        const bool sbp = fused && opt_img_sbp && sub == 1;
        Levels A = levels_in(d_adj, sbp ? 4 : 3);
        init_levels(A, true, dadj);
        if (sbp) {
            const int seg0 = seg_loaded;
            run_loop(graph_adj, loop_key(2, dadj, dsrc, seg0), [&] {
                Levels Ac = A;
                img_out[0] = img_out[1] = false;
                for (int it = 0; it < nt - 1; it++) {
                    const int j = nt - 1 - it;
                    load_segment(j / seg_K, dsrc);
                    stat_image_calls++;
                    step(Ac, it, true, dadj, nullptr, j, true);
                }
                // snapshot 0: the last level and the one before it
                load_segment(0, dsrc);
                StepArgs ia = base_args();
                sbp_image_args(ia, 0);
                ia.spb = std::max(1, std::min(opt_spb_img > 0 ? opt_spb_img : 4, ns));
                ia.ngroups = (ns + ia.spb - 1) / ia.spb;
                ia.aprev = Ac.P();
                ia.curr = Ac.C();
                ia.next = Ac.C();  // weight 0
                // fused: the stiffness terms of snapshot 0 came with the last step
                if (image_fused()) ia.img_deriv = 0;
                stat_image_calls++;
                run_img(ia, nt - 1, true);
                img_join();
            });
        } else if (fused) {
            // step it takes the adjoint from level it to it+1; the pair (it, it+1) images
            // snapshot (nt-2-it)/sub when that divides (the reference images at its
            // iteration it+1, after the step, with the derivatives of level it)
            const int seg0 = seg_loaded;
            run_loop(graph_adj, loop_key(1, dadj, dsrc, seg0), [&] {
                Levels Ac = A;
                img_out[0] = img_out[1] = false;
                for (int it = 0; it < nt - 1; it++) {
                    int idx = -1;
                    if ((nt - 2 - it) % sub == 0) {
                        idx = (nt - 2 - it) / sub;
                        load_segment(idx / seg_K, dsrc);
                        stat_image_calls++;
                    }
                    step(Ac, it, true, dadj, nullptr, idx);
                }
                img_join();
            });
        } else {
            for (int it = 0; it < nt - 1; it++) {
                if (it != 0 && (nt - 1 - it) % sub == 0) image(A, (nt - 1 - it) / sub, dsrc);
                step(A, it, true, dadj, nullptr);
            }
            image(A, 0, dsrc);
        }
        d_gout.reserve((size_t)np * nxz * sizeof(float));
        float* out = d_gout.as<float>();
        launch_sum_shots(d_gshot.as<float>(), (long long)np * nxz, nslab, (long long)np * nxz, out, 0, st);
        if (opt_timing) {
            HIPCHECK(hipEventRecord(ev1, st));
            HIPCHECK(hipEventSynchronize(ev1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
            last_adj_ms = ms;
            (void)hipEventDestroy(ev0);
            (void)hipEventDestroy(ev1);
        }
        // [np][nx][ldz] -> [np][nx][nz]
        HIPCHECK(hipMemcpy2DAsync(grad, (size_t)g.nz * sizeof(float), out, (size_t)ldz * sizeof(float),
                                  (size_t)g.nz * sizeof(float), (size_t)np * g.nx,
                                  grad_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipGetLastError());
    }

    double dot_f64(const float* x, const float* y, const float* w, long long n)
    {
        const int nb = kDotBlocks;
        d_partial.reserve(nb * sizeof(double));
        d_scalar.reserve(sizeof(double));
        launch_dot_f64(x, y, w, n, d_partial.as<double>(), nb, st);
        launch_reduce_f64(d_partial.as<double>(), nb, d_scalar.as<double>(), st);
        double out = 0;
        HIPCHECK(hipMemcpyAsync(&out, d_scalar.p, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return out;
    }
};

}  // namespace fwi2d

using fwi2d::Engine;
struct fwi2d_engine {
    Engine* impl;
};

// Device memory of the allocation-free entry point fwi2d_dot_f64_noalloc: the block
// partials and the result of one dot product, (kDotBlocks + 1) doubles = 8 KB on every
// device the library runs on, held until the process ends.  Host code reaches it only
// through the address that hipGetSymbolAddress returns for the current device
// (noalloc_scratch); the array's name on the host is not a device address.
static __device__ __align__(256) double scratch[fwi2d::kDotBlocks + 1];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), resolved once per
// device on its first use
double* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[device] = static_cast<double*>(p);
    }
    return g_scratch[device];
}
}  // namespace

template <class F>
static int guarded(F&& f)
{
    try {
        f();
        fwi2d::g_last_error.clear();
        return 0;
    } catch (const std::exception& ex) {
        fwi2d::g_last_error = ex.what();
        return 1;
    } catch (...) {
        fwi2d::g_last_error = "unknown error";
        return 1;
    }
}

extern "C" {

const char* fwi2d_last_error(void) { return fwi2d::g_last_error.c_str(); }

int fwi2d_engine_create(const fwi2d_grid* grid, int device, fwi2d_engine** out)
{
    return guarded([&] {
        if (!grid || !out) throw std::runtime_error("null argument");
        *out = nullptr;
        Engine* e = new Engine(*grid, device);
        *out = new fwi2d_engine{e};
    });
}

int fwi2d_engine_set_model(fwi2d_engine* e, const float* model)
{
    return guarded([&] {
        if (!e || !model) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->set_model(model);
    });
}

int fwi2d_engine_set_shots(fwi2d_engine* e, int nshots, const double* sxz, const int* nrec, const double* rxz,
                           double ox, double oz)
{
    return guarded([&] {
        if (!e || !sxz || !nrec) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->set_shots(nshots, sxz, nrec, rxz, ox, oz);
    });
}

int fwi2d_engine_set_acquisition(fwi2d_engine* e, int nshots, const double* sxz, const int* nrec, const double* rxz,
                                 double ox, double oz, int moment_tensor, double gauge_length, const double* rdip)
{
    return guarded([&] {
        if (!e || !sxz || !nrec) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->set_shots(nshots, sxz, nrec, rxz, ox, oz, moment_tensor != 0, gauge_length, rdip);
    });
}

int fwi2d_engine_forward(fwi2d_engine* e, const float* src, float* rcv, int keep_checkpoints, int on_device)
{
    return guarded([&] {
        if (!e || !src) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->forward(src, rcv, keep_checkpoints != 0, on_device != 0, on_device != 0);
    });
}

int fwi2d_engine_adjoint(fwi2d_engine* e, const float* adj, float* grad, int on_device)
{
    return guarded([&] {
        if (!e || !adj || !grad) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        e->impl->adjoint(adj, grad, on_device != 0, on_device != 0);
    });
}

int fwi2d_engine_dot_f64(fwi2d_engine* e, const float* a, const float* b, const float* w, long long n, double* out)
{
    return guarded([&] {
        if (!e || !a || !out) throw std::runtime_error("null argument");
        HIPCHECK(hipSetDevice(e->impl->device));
        *out = e->impl->dot_f64(a, b, w, n);
    });
}

int fwi2d_dot_f64_noalloc(int device, const float* a, const float* b, const float* w, long long n, double* out)
{
    int prev = -1;
    const int rc = guarded([&] {
        if (!a || !out) throw std::runtime_error("null argument");
        if (n < 0 || n > FWI2D_DOT_NOALLOC_MAX_N)
            throw std::runtime_error("fwi2d_dot_f64_noalloc: n = " + std::to_string(n) +
                                     " outside [0, FWI2D_DOT_NOALLOC_MAX_N]");
        int ndev = 0;
        HIPCHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw std::runtime_error("no HIP device " + std::to_string(device));
        HIPCHECK(hipGetDevice(&prev));
        HIPCHECK(hipSetDevice(device));
        double* part = noalloc_scratch(device);
        // the null stream: ordered after earlier work on the library's (blocking) streams
        fwi2d::launch_dot_f64(a, b, w, n, part, fwi2d::kDotBlocks, 0);
        fwi2d::launch_reduce_f64(part, fwi2d::kDotBlocks, part + fwi2d::kDotBlocks, 0);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpy(out, part + fwi2d::kDotBlocks, sizeof(double), hipMemcpyDeviceToHost));
    });
    if (prev >= 0) (void)hipSetDevice(prev);
    return rc;
}

int fwi2d_engine_set_option(fwi2d_engine* e, const char* key, long long value)
{
    return guarded([&] {
        if (!e || !key) throw std::runtime_error("null argument");
        Engine* x = e->impl;
        std::string k(key);
        if (k == "tile_x") {
            x->opt_tile_x = (int)value;
            x->set_tile((int)value);
        } else if (k == "segment") {
            if (value < 0) throw std::runtime_error("segment must be >= 0");
            x->opt_segment = (int)value;
            x->ckpt_epoch = -1;
        } else if (k == "checkpoint_budget") {
            if (value < 0) throw std::runtime_error("checkpoint_budget must be >= 0");
            x->opt_ckpt_budget_mb = value;
            x->ckpt_epoch = -1;
        } else if (k == "lag") {
            x->opt_lag = value ? 1 : 0;
        } else if (k == "timing") {
            x->opt_timing = value ? 1 : 0;
        } else if (k == "graph") {
            x->opt_graph = value ? 1 : 0;  // captured loops stay cached, keyed on their configuration
        } else if (k == "shots_per_block") {
            if (value < 1 || value > 64) throw std::runtime_error("shots_per_block must be in [1, 64]");
            x->opt_spb = (int)value;
        } else if (k == "stiffness_gradients") {
            x->opt_stiff = value ? 1 : 0;
        } else if (k == "checkpoint_format") {
            if (value < 0 || value > 2) throw std::runtime_error("checkpoint_format must be 0 (fp32), 1 (bf16) or 2 (fp16)");
            x->opt_bf16 = (int)value;
            x->ckpt_epoch = -1;
        } else if (k == "edge_stream") {
            if (value < 0 || value > 2) throw std::runtime_error("edge_stream must be 0, 1 or 2");
            x->opt_edge_stream = (int)value;
        } else if (k == "autotune") {
            x->opt_autotune = value ? 1 : 0;
            x->tuned = false;
        } else if (k == "shots_per_block_img") {
            if (value < 0 || value > 64) throw std::runtime_error("shots_per_block_img must be in [0, 64]");
            x->opt_spb_img = (int)value;
        } else if (k == "shots_per_block_edge") {
            if (value < 1 || value > 64) throw std::runtime_error("shots_per_block_edge must be in [1, 64]");
            x->opt_spb_edge = (int)value;
        } else if (k == "band_rows") {
            if (value != 8 && value != 12 && value != 16) throw std::runtime_error("band_rows must be 8, 12 or 16");
            x->opt_band_rows = (int)value;
            x->build_boxes();
        } else if (k == "shots_per_block_edge_img") {
            if (value < 0 || value > 64) throw std::runtime_error("shots_per_block_edge_img must be in [0, 64]");
            x->opt_spb_edge_img = (int)value;
        } else if (k == "image_fused") {
            if (value < -1 || value > 1) throw std::runtime_error("image_fused must be -1, 0 or 1");
            x->opt_image_fused = (int)value;
        } else if (k == "image_sbp") {
            x->opt_img_sbp = value ? 1 : 0;
        } else if (k == "image_stream") {
            x->opt_img_stream = value ? 1 : 0;
        } else if (k == "image_buffers") {
            if (value != 1 && value != 2) throw std::runtime_error("image_buffers must be 1 or 2");
            x->opt_img_buffers = (int)value;
        } else {
            throw std::runtime_error("unknown option " + k);
        }
    });
}

long long fwi2d_engine_get_stat(fwi2d_engine* e, const char* key)
{
    if (!e || !key) return -1;
    Engine* x = e->impl;
    std::string k(key);
    if (k == "steps") return x->stat_steps;
    if (k == "recompute_steps") return x->stat_recompute_steps;
    if (k == "image_calls") return x->stat_image_calls;
    if (k == "segment") return x->seg_K;
    if (k == "nseg") return x->nseg;
    if (k == "tile_x") return x->tile_x;
    if (k == "shots_per_block") return x->opt_spb;
    if (k == "shots_per_block_edge") return x->opt_spb_edge;
    if (k == "shots_per_block_edge_img") return x->opt_spb_edge_img;
    if (k == "band_rows") return x->opt_band_rows;
    if (k == "shots_per_block_img") return x->opt_spb_img;
    if (k == "edge_stream") return x->opt_edge_stream;
    if (k == "image_fused") return x->image_fused() ? 1 : 0;
    if (k == "image_buffers") return x->opt_img_buffers;
    if (k == "image_stream") return x->opt_img_stream;
    if (k == "image_sbp") return x->opt_img_sbp;
    if (k == "ntiles") return x->ntiles();
    if (k == "interior_tiles") return (long long)x->list_core.size();
    if (k == "graph_captures") return x->graph_captures;
    if (k == "graph_replays") return x->graph_replays;
    if (k == "nrec") return x->nr;
    if (k == "nshots") return x->ns;
    if (k == "fwd_loop_ns") return (long long)(x->last_fwd_loop_ms * 1e6);
    if (k == "fwd_loop_steps") return x->last_fwd_loop_steps;
    if (k == "adj_ns") return (long long)(x->last_adj_ms * 1e6);
    return -1;
}

const char* fwi2d_engine_tune_report(fwi2d_engine* e)
{
    if (!e) return "";
    return e->impl->tune_report.c_str();
}

void fwi2d_engine_destroy(fwi2d_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->impl->device);
    delete e->impl;
    delete e;
}

}  // extern "C"

namespace fwi2d {

void engine_forward_traces(fwi2d_engine* e, const float* src, bool keep, const TraceMatrix& m, float* out)
{
    if (!e || !src || !out) throw std::runtime_error("null argument");
    HIPCHECK(hipSetDevice(e->impl->device));
    e->impl->forward_traces(src, keep, m, out);
}

void engine_adjoint_traces(fwi2d_engine* e, const float* data, const TraceMatrix& m, float* grad)
{
    if (!e || !data || !grad) throw std::runtime_error("null argument");
    HIPCHECK(hipSetDevice(e->impl->device));
    e->impl->adjoint_traces(data, m, grad);
}

}  // namespace fwi2d
