// Kernel argument blocks and launchers of the elastic propagation engine (gfx950).
//
// Device layout: every field is [x][ldz] with z contiguous and ldz = nz rounded up to a
// multiple of 32 (128-byte columns, so that a 64-cell column segment of a tile is two
// whole cache lines); each thread moves 4 consecutive z cells (a "quad") with one
// 128-bit access; the padding cells are zero and never read as stencil data.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fwi2d {

constexpr int kTileZ = 64;     // z extent of a tile: 16 quads
constexpr int kThreads = 256;  // 4 wavefronts per workgroup

// One 3x3 footprint of a point injector or extractor, starting at (xi0, zi0), with
// separable weights wx[i] wz[j]; it feeds (or is read into) trace sample offset `trace`
// for the x field component and `tz` for the z component.  A delta_m3 point is one
// footprint (trace and tz the two components of one trace), a ddelta_m3 moment tensor
// two (the x- and z-derivative parts, reading the mxx/mxz and mxz/mzz traces), a
// dipole_m3 DAS channel two (the gauge ends, with opposite signs).
struct PointEntry {
    int xi0, zi0;
    float wx[3], wz[3];
    long long trace, tz;
    int shot;
    int pad;
};

// A cell touched by point injectors: shot and grid position; its contributions are
// listed in a CSR (trace offsets of the x and z field components, weight) in injector
// order.
struct InjCell {
    int shot, gx, gz, pad;
};

struct StepArgs {
    int nx, nz, ldz, ntx, ntz, nshots, nt, it, taper_prev;
    int spb, ngroups;      // shots per workgroup, shot groups (ceil(nshots / spb))
    float rdx, rdz, rdx2, rdz2, dt, dt2, sat_x, sat_z;
    float c1x, c2x, c1z, c2z;  // interior first-derivative coefficients divided by dx, dz
    int bc_top, bc_bottom, bc_left, bc_right;
    const float *lam, *mu, *rho, *c13, *eps;
    const float* irho;     // dt^2 / rho
    const float *mux, *lamx;  // VTI: (1 + 2 eps) mu, (1 + 2 eps) lambda
    const float *div_tb;   // [2 sides][2 comps][nx] absorbing divisors, top then bottom
    const float *div_lr;   // [2 sides][2 comps][nz]
    const float *taper_x;  // [nx]
    const float *taper_z;  // [ldz]
    const float* prev;     // shot 0 base of [2][nx][ldz]; shot stride below
    const float* curr;
    float* next;
    long long shot_stride;
    // imaging (k_img after an adjoint step, or fused into it): the derivatives of the
    // step's current level and its new level image forward snapshot idx
    // (compute_gradients, we_op.cpp:505-562)
    int img_mode;              // 0 interior idx, 1 first (idx 0), 2 last
    float wlm, wrho;           // lambda/mu and rho weights of this snapshot
    const float *f0, *fm, *fp; // forward snapshots idx, idx-1, idx+1 (shot 0 base)
    long long fshot_stride;
    float* grad;               // raw gradient slabs [shot][np][nx][ldz]; a workgroup adds its
    long long gslab_stride;    // shots' sum to the slab of its first shot
    const float *ge, *gv2, *gv3;  // VTI: 1 + 2 eps, d c13 / d lambda, d c13 / d mu
    int stiff;                 // VTI: image c11, c13, c33, c55, rho instead of lambda, mu, rho
    // density term by parts in time (k_img, one snapshot per step): f0 is snapshot j, its
    // derivatives meet those of aprev; the density term is f0 . (wn next + wc curr + wp
    // aprev) with the second-difference weights of the neighbouring snapshots
    const float* aprev;
    float wn, wc, wp;
    int img_deriv;             // 0: snapshot j has no stiffness terms (only density terms)
};

struct ImageArgs {
    int nx, nz, ldz, ntx, ntz, nshots, mode;  // mode 0 interior, 1 first (idx 0), 2 last
    float rdx, rdz, wlm, wrho;                // weights: lambda/mu (w*dtg or w*dtg/2), rho (w/dtg)
    const float *lam, *mu, *c13, *eps;
    const float* adj;        // current adjoint level, shot 0 base
    const float* adj_prev;   // previous adjoint level (its derivatives are imaged)
    long long shot_stride;
    const float* f0;         // forward snapshot idx   [shot][2][nx][ldz], shot stride fshot_stride
    const float* fm;         // forward snapshot idx-1 (mode 0, 2)
    const float* fp;         // forward snapshot idx+1 (mode 0, 1)
    long long fshot_stride;
    float* grad;             // per-shot raw gradients [shot][ncomp][nx][ldz]
    int ncomp;
    long long gshot_stride;
    int lag;                 // 1: derivatives of adj_prev (reference), 0: of adj
};

void upload_constants(hipStream_t s);

int tiles_x(int nx, int tile_x);
int tiles_z(int ldz);
void tile_range_x(int k, int ntile, int T, int n, int& s, int& e);
void tile_range_z(int k, int ntile, int nz, int ldz, int& s, int& e);

// Box shapes of the step decomposition: core boxes (TX x 64, every stencil interior) and
// edge-band boxes (120 x 8 for 8-row bands and 64 x 16 for taller ones along the top and
// bottom, 8 x 128 along the sides).  The 8-row boxes are 120 wide so that their state
// window is 128 columns: two window quads and two product quads per thread.
enum class BoxClass { core16, core32, band_z, band_x, band_z8 };
constexpr int kBandZ8Width = 120;
constexpr int kCoreX0 = 6;   // core: x in [6, nx - 6), z in [8, core_z_end(nz))
constexpr int kCoreZ0 = 8;
inline int core_z_end(int nz) { return ((nz - 8) / 4) * 4; }

// n boxes listed in list (indices into boxes: {sx, ex, sz, ez}) of one class
void launch_step(const StepArgs& a, bool vti, BoxClass c, const int4* boxes, const int* list, int n, bool image,
                 hipStream_t s);
void launch_init_prev(float* prev, long long shot_stride, long long nxz, int ldz, int nshots, const int* shot_ptr,
                      const PointEntry* ents, const float* trace, hipStream_t s);
void launch_scale_prev(float* prev, const float* rho, float half_dt2, long long nxz, int nshots, hipStream_t s);
// Density-term share of the point sources of a fused imaging step (by parts in time,
// on = 0: no imaging): wn f0 . (injected x, z) into the shot's gradient slab at rho_off,
// the offset of the density component in a slab.
struct InjImage {
    int on;
    float wn;
    const float* f0;
    long long fshot_stride;
    float* grad;
    long long gslab_stride, rho_off;
};
void launch_inject_cells(const InjCell* cells, int ncell, const int* ptr, const long long* ctx, const long long* ctz,
                         const float* cw, const float* trace, int it, float* next, long long shot_stride,
                         long long nxz, int nx, int nz, int ldz, const float* irho, const float* taper_x,
                         const float* taper_z, const float* div_tb, const float* div_lr, const InjImage& im,
                         hipStream_t s);
// receiver r sums the footprints ents[rptr[r]] .. ents[rptr[r+1]-1] into its traces
void launch_extract(const float* state, long long shot_stride, long long nxz, int ldz, const PointEntry* ents,
                    const int* rptr, int nrec, float* trace, int it, hipStream_t s);
void launch_image(const ImageArgs& a, bool vti, int tile_x, hipStream_t s);
// imaging condition after an adjoint step over the tile_x x 64 tile grid (a.ntx x a.ntz
// tiles, a.ngroups groups of a.spb shots); a.curr lagged adjoint level, a.next the new
// one (sbp: the density term by parts, a.aprev the lagged level, see StepArgs); nbuf 2:
// the next shot's windows load into a second LDS buffer during this shot
void launch_img(const StepArgs& a, bool vti, int tile_x, int nbuf, bool sbp, hipStream_t s);
void launch_sum_shots(const float* per_shot, long long shot_stride, int nshots, long long n, float* out,
                      int accumulate, hipStream_t s);
void launch_fill(float* p, long long n, float v, hipStream_t s);
// out [ntr][nrow] = per-trace row-compressed matrix (ptr [nrow + 1], col, w) times in [ntr][ncol]
void launch_trace_matvec(const int* ptr, const int* col, const double* w, int nrow, int ncol, const float* in,
                         float* out, long long ntr, hipStream_t s);
void launch_copy(float* dst, const float* src, long long n, hipStream_t s);
// dipole_to_strain of nr DAS channels in place on traces [2][nr][nt] (cs: cos, sin of
// each channel's dip; forward or adjoint)
void launch_das_strain(float* tr, const double* cs, int nr, int nt, double gl, bool adjoint, hipStream_t s);
// per-shot max-abs scaled 16-bit packing of a time level, FP16 (f16) or BF16 (amax:
// nshots words of scratch)
void launch_pack16(bool f16, const float* u, unsigned short* q, float* scale, unsigned* amax, long long per_shot,
                   int nshots, hipStream_t s);
void launch_unpack16(bool f16, const unsigned short* q, const float* scale, float* u, long long per_shot, int nshots,
                     hipStream_t s);
void launch_dot_f64(const float* a, const float* b, const float* w, long long n, double* partial,
                    int nblocks, hipStream_t s);
void launch_reduce_f64(const double* partial, int n, double* out, hipStream_t s);

// ---------------------------------------------------------------------------
// Acoustic operator (nl_we_op_a, we_op.cpp:1863-2135): one field per shot, [nshots][nx][ldz]
// ---------------------------------------------------------------------------
constexpr int kAcTX = 32, kAcTZ = 32;  // step and imaging tiles (x, z)

struct AcousticArgs {
    int nx, nz, ldz, nshots, ntx, ntz, taper_prev;
    int spb;                     // k_astep: shots per workgroup (consecutive shots of one tile)
    int itx0, itx1, itz0, itz1;  // k_astep: the tiles [itx0, itx1) x [itz0, itz1) hold interior cells only
    int bc[4];                   // top, bottom, left, right: 1 Dirichlet, 2 absorbing, 3 Neumann
    float rdx2, rdz2;            // 1 / dx^2, 1 / dz^2
    float dir_x[4], dir_z[4];    // Dirichlet SAT S^T terms SC[k] / (H[k] h^2)
    float pen_x, pen_z;          // Dirichlet penalty 1 / (a H0 h^2)
    float neu_x, neu_z;          // Neumann / absorbing 1 / (H0 h^2)
    float abs_x, abs_z;          // absorbing 1 / (2 dt H0 h)
    const float *K, *ri, *kdt2, *div, *taper_x, *taper_z;  // [nx][ldz] (tapers per axis)
    const float* prev;
    const float* curr;
    float* next;
    long long shot_stride;       // nx * ldz
    // imaging (k_agrad): adjoint level curr against snapshots f0 (idx), fm, fp
    const float *f0, *fm, *fp;
    long long fshot_stride;
    float* grad;                 // per shot-group slabs [groups][2][nx][ldz] (groups of spb shots)
    long long gslab_stride;
    int mode;                    // 0 interior snapshot, 1 the first, 2 the last
    int img;                     // k_astep: image snapshot f0 with the adjoint level curr as well
    float wk;                    // K term weight w / dtg
    float wr;                    // 1/rho term weight w dtg (halved for the first snapshot)
    float rdx, rdz;              // 1 / dx, 1 / dz
};
// one time step of every shot: next from prev and curr (no point sources)
void launch_astep(const AcousticArgs& a, hipStream_t s);
// imaging condition of one snapshot for every shot, into the per-shot slabs
void launch_agrad(const AcousticArgs& a, hipStream_t s);
// point sources: next[cell] += scale[cell] * sum_k w[k] trace[off[k] + it] / div * taper
// (scale = dt^2 K; with init: prev = 1/2 dt^2 K sum_k ..., no divisor or taper)
void launch_ainject(const InjCell* cells, int ncell, const int* ptr, const long long* off, const float* w,
                    const float* trace, int it, float* u, long long shot_stride, int ldz, const float* kdt2,
                    const float* div, const float* taper_x, const float* taper_z, bool init, hipStream_t s);
// receivers: trace[r][it] = sum over the 9 footprint cells of w * u (shot of the receiver)
void launch_aextract(const float* u, long long shot_stride, int ldz, const PointEntry* ents, int nrec, float* trace,
                     int it, hipStream_t s);

}  // namespace fwi2d
