// Device kernels of the elastic SBP-SAT propagation engine, written for gfx950.
//
// Layout: every field is [x][ldz] with z contiguous (index ix*ldz + iz), the reference's
// [x][z] order with z padded to a multiple of 32 cells (128-byte columns).  Workgroups
// of 256 threads own a box of cells and a group of shots; a thread moves 4 consecutive
// z cells (a quad) with 128-bit accesses.  The time step runs over a box decomposition
// (interior core boxes and edge bands, built by the engine); the imaging kernel runs
// over the tile grid below.
//
// Tile grid: tiles start at multiples of the tile size except the last one in each
// direction, which is anchored to end at the grid edge (start = n - T).  Each tile
// writes the cells in [start, next tile start) (the last tile writes to the edge), so
// writes never overlap and the SBP closure rows near the far edge are always written by
// the anchored tile.
#include "kernels.h"
#include "sbp_tables.h"
#include <hip/hip_fp16.h>

namespace fwi2d {

__constant__ float cD1I[2];
__constant__ float cSC[4];
__constant__ float cD2I[10];

void upload_constants(hipStream_t s)
{
    float d1i[2], sc[4], d2i[10];
    d1i[0] = (float)kD1Interior[0];
    d1i[1] = (float)kD1Interior[1];
    for (int i = 0; i < 4; i++) sc[i] = (float)kBoundaryDeriv[i];
    for (int i = 0; i < 10; i++) d2i[i] = (float)kD2Interior[i];
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD1I), d1i, sizeof(d1i), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cSC), sc, sizeof(sc), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD2I), d2i, sizeof(d2i), 0, hipMemcpyHostToDevice, s);
    (void)hipStreamSynchronize(s);
}


// ---------------------------------------------------------------------------
// Tile geometry.  x: tiles of TX columns; z: tiles of 64 cells aligned on quads.  The
// last tile of each axis is anchored at the far edge so that the SBP closure rows there
// are always written by one tile; write ranges partition the grid.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void tile_range_x_hd(int k, int ntile, int T, int n, int& s, int& e)
{
    if (ntile == 1) { s = 0; e = n; return; }
    if (k == ntile - 1) { s = n - T; e = n; return; }
    s = k * T;
    e = (s + T < n - T) ? s + T : n - T;
}

__host__ __device__ __forceinline__ void tile_range_z_hd(int k, int ntile, int nz, int ldz, int& s, int& e)
{
    const int last = ldz - kTileZ;  // quad-aligned start of the anchored tile
    if (ntile == 1) { s = 0; e = nz; return; }
    if (k == ntile - 1) { s = last; e = nz; return; }
    s = k * kTileZ;
    e = (s + kTileZ < last) ? s + kTileZ : last;
}

int tiles_x(int nx, int tile_x) { return (nx + tile_x - 1) / tile_x; }
int tiles_z(int ldz) { return (ldz + kTileZ - 1) / kTileZ; }
void tile_range_x(int k, int ntile, int T, int n, int& s, int& e) { tile_range_x_hd(k, ntile, T, n, s, e); }
void tile_range_z(int k, int ntile, int nz, int ldz, int& s, int& e) { tile_range_z_hd(k, ntile, nz, ldz, s, e); }

// first cell of a tile's window with halo H such that SBP closures at the near edge
// (rows < 6 for D2, < 4 for D1) always see the edge itself
__device__ __forceinline__ int window_start(int s, int H)
{
    return (s >= H && s < H + 2) ? 0 : s - H;
}

// Range [lo, hi) of first-derivative products a tile writing [s, e) of n cells reads:
// its cells +/- 2, widened to the 6 closure inputs of D1 rows at either grid edge.
__host__ __device__ __forceinline__ void product_range(int s, int e, int n, int& lo, int& hi)
{
    lo = s - 2 > 0 ? s - 2 : 0;
    hi = e + 2 < n ? e + 2 : n;
    if (s < 4) { lo = 0; hi = hi > 6 ? hi : 6; }
    if (e > n - 4) { lo = lo < n - 6 ? lo : n - 6; hi = n; }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous run
// of (shot, tile) work items so that neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = bid & 7, local = bid >> 3;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// every quad access is 16-byte aligned (quads start at multiples of 4 cells of 16-byte
// aligned arrays); saying so lets LDS quads load with one 128-bit read
__device__ __forceinline__ float4 ld4(const float* p)
{
    return *reinterpret_cast<const float4*>(__builtin_assume_aligned(p, 16));
}
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(__builtin_assume_aligned(p, 16)) = v; }
__device__ __forceinline__ void unpack(float4 v, float* d)
{
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
    d[3] = v.w;
}

// ---------------------------------------------------------------------------
// SBP stencils on LDS lines.  p points at the element of global index g (0 <= g < n);
// s is the LDS stride along the differentiated axis.  Results are not divided by h.
// Closure rows are evaluated out of line so that the interior path stays small.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float d1_int(const float* p, int s)
{
    return cD1I[0] * (p[s] - p[-s]) + cD1I[1] * (p[2 * s] - p[-2 * s]);
}

// Closure row R of D1 from the edge point q along stride st, with the coefficients as
// immediates (zero ones skipped).
template <int R>
__device__ __forceinline__ float d1_closure_row(const float* q, int st)
{
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < 6; j++)
        if (kD1Closure[R][j] != 0.0) v += (float)kD1Closure[R][j] * q[j * st];
    return v;
}

// D1 closure row of point g of n: rows 0..3 at the near edge, mirrored with opposite sign
// at the far edge.  The row is dispatched to a compile-time instance (a uniform branch
// where the row is the same for a whole wavefront).
__device__ __forceinline__ float d1_closure(const float* p, int s, int g, int n)
{
    const bool near = g < 4;
    const int r = near ? g : n - 1 - g;
    const float* q = near ? p - g * s : p + r * s;
    const int st = near ? s : -s;
    float v;
    switch (r) {
    case 0: v = d1_closure_row<0>(q, st); break;
    case 1: v = d1_closure_row<1>(q, st); break;
    case 2: v = d1_closure_row<2>(q, st); break;
    default: v = d1_closure_row<3>(q, st); break;
    }
    return near ? v : -v;
}

template <bool IN>
__device__ __forceinline__ float d1(const float* p, int s, int g, int n)
{
    if (IN || (g >= 4 && g < n - 4)) return d1_int(p, s);
    return d1_closure(p, s, g, n);
}

// interior weights of D2(c) at one point: u[-2..2] coefficients (Dzz_var/Dxx_var interior)
struct W5 {
    float m2, m1, c, p1, p2;
};

__device__ __forceinline__ W5 d2_weights(const float* pc, int s)
{
    const float fm2 = pc[-2 * s], fm1 = pc[-s], f0 = pc[0], fp1 = pc[s], fp2 = pc[2 * s];
    W5 w;
    w.c = cD2I[0] * f0 + cD2I[1] * (fm1 + fp1) + cD2I[2] * (fm2 + fp2);
    w.m1 = cD2I[3] * fm2 + cD2I[4] * fm1 + cD2I[5] * f0 + cD2I[6] * fp1;
    w.p1 = cD2I[3] * fp2 + cD2I[4] * fp1 + cD2I[5] * f0 + cD2I[6] * fm1;
    w.m2 = cD2I[7] * fm2 + cD2I[8] * fm1 + cD2I[9] * f0;
    w.p2 = cD2I[7] * fp2 + cD2I[8] * fp1 + cD2I[9] * f0;
    return w;
}

// D2 weights of the 4 cells of a quad along x and z (coefficient mu); they depend on the
// model only, so a workgroup forms them once and reuses them for all its shots
struct QuadW {
    W5 x[4], z[4];
};

__device__ __forceinline__ W5 d2_scaled(W5 w, float f)
{
    w.m2 *= f;
    w.m1 *= f;
    w.c *= f;
    w.p1 *= f;
    w.p2 *= f;
    return w;
}

__device__ __forceinline__ float d2_apply(const W5& w, const float* pu, int s)
{
    return w.c * pu[0] + w.m1 * pu[-s] + w.p1 * pu[s] + w.m2 * pu[-2 * s] + w.p2 * pu[2 * s];
}

// Closure rows of D2(c): row R of the 6 x 8 x 8 table applied to one or two fields
// sharing the coefficient, from the edge point along stride st.  The table's non-zeros
// are compile-time immediates; the weights of the 8 field points are formed first.
constexpr int kD2NZ = sizeof(kD2ClosureNZ) / sizeof(kD2ClosureNZ[0]);
constexpr int d2_row_begin(int r)
{
    int i = 0;
    while (i < kD2NZ && kD2ClosureNZ[i].row < r) i++;
    return i;
}

template <int R, bool TWO>
__device__ __forceinline__ void d2_closure_row(const float* qu, const float* qv, const float* qc, int st, float& ou,
                                               float& ov)
{
    float val[8];
#pragma unroll
    for (int j = 0; j < 8; j++) val[j] = 0.f;
    constexpr int b = d2_row_begin(R), e = d2_row_begin(R + 1);
#pragma unroll
    for (int i = b; i < e; i++)
        val[kD2ClosureNZ[i].point] +=
            (float)(kD2ClosureNZ[i].num / kD2ClosureNZ[i].den) * qc[kD2ClosureNZ[i].coef * st];
    float cu = 0.f, cv = 0.f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        cu += val[j] * qu[j * st];
        if (TWO) cv += val[j] * qv[j * st];
    }
    ou = cu;
    ov = cv;
}

// closure rows of D2(c) for one field (TWO = false) or two fields sharing the coefficient
template <bool TWO>
__device__ __forceinline__ void d2_closure_any(const float* pu, const float* pv, const float* pc, int s, int g, int n,
                                               float& ou, float& ov)
{
    const bool top = g < 6;
    const int r = top ? g : n - 1 - g;
    const int st = top ? s : -s;
    const int o = top ? -g * s : r * s;
    const float *qu = pu + o, *qv = pv + o, *qc = pc + o;
    switch (r) {
    case 0: d2_closure_row<0, TWO>(qu, qv, qc, st, ou, ov); break;
    case 1: d2_closure_row<1, TWO>(qu, qv, qc, st, ou, ov); break;
    case 2: d2_closure_row<2, TWO>(qu, qv, qc, st, ou, ov); break;
    case 3: d2_closure_row<3, TWO>(qu, qv, qc, st, ou, ov); break;
    case 4: d2_closure_row<4, TWO>(qu, qv, qc, st, ou, ov); break;
    default: d2_closure_row<5, TWO>(qu, qv, qc, st, ou, ov); break;
    }
}

__device__ __forceinline__ float d2_closure(const float* pu, const float* pc, int s, int g, int n)
{
    float ou, ov;
    d2_closure_any<false>(pu, pu, pc, s, g, n, ou, ov);
    return ou;
}

__device__ __forceinline__ void d2_closure2(const float* pu, const float* pv, const float* pc, int s, int g, int n,
                                            float& ou, float& ov)
{
    d2_closure_any<true>(pu, pv, pc, s, g, n, ou, ov);
}

// one-sided outward boundary derivative sum_k SC[k] u[k*s] (times h)
__device__ __forceinline__ float bnd_deriv(const float* p, int s)
{
    return cSC[0] * p[0] + cSC[1] * p[s] + cSC[2] * p[2 * s] + cSC[3] * p[3 * s];
}

// sum_j B[0][j] u[j*s]: first closure row (times h), used by the normal complement terms
__device__ __forceinline__ float d1_row0(const float* p, int s) { return d1_closure_row<0>(p, s); }

// ---------------------------------------------------------------------------
// Fused time step
//
// A workgroup owns a 64 (z) x TX (x) tile and advances a group of shots through it one
// after the other, so that the coefficient window and the per-cell factors are read
// once per tile, not once per shot.  Per shot: the state window (halo 4 in x, 8 in z)
// is staged in LDS, the products of first derivatives are formed on the tile +/- 2 (x)
// and +/- 4 (z), then each thread computes the new level of its quads (4 consecutive z
// cells).  The x and z parts of every stencil choose between the vectorised interior
// path and the closure path independently, so only quads that really touch a closure
// row take it.  The next shot's window is loaded into registers before the current
// shot's arithmetic, which hides its latency.
// ---------------------------------------------------------------------------
// Thin z bands (TZ <= 16) map lanes row-major: a wavefront holds one quad row of 64
// consecutive columns, so the z closure row of every lane of a wave is the same and the
// closure coefficients are read with scalar loads.  Their LDS columns then hold an odd
// number of quads, which keeps the column-strided 16-byte reads free of bank conflicts.
// 8-row bands (TZ = 8) hold compact windows: 16 state rows and 12 product rows, starting
// at row 0 for the top band; elsewhere the state rows start 8 and the product rows 4
// above the box (the interior product quads read 4 state rows above themselves, the
// bottom closure rows the 8 rows above the last one).
template <int TX, int TZ_>
struct StepShape {
    static constexpr int TZ = TZ_;
    static constexpr bool ROWMAJ = TZ <= 16;
    static constexpr int BX = ROWMAJ ? (TX + 63) / 64 * 64 : TX;  // lanes per quad row (row-major)
    static constexpr bool COMPACT = TZ == 8;
    static constexpr int HX = 4, HZ = 8;
    static constexpr int WX = TX + 2 * HX;                     // state window columns
    static constexpr int WQ = COMPACT ? 4 : (TZ + 2 * HZ) / 4; // state window quads per column
    static constexpr int WZ = 4 * (ROWMAJ ? (WQ | 1) : WQ);    // LDS column stride of the window
    static constexpr int IX = TX + 4;                          // product columns
    static constexpr int PQ = COMPACT ? 3 : (TZ + 8) / 4;      // product quads: rows [sz-4, sz+TZ+4)
    static constexpr int PZ = 4 * (ROWMAJ ? (PQ | 1) : PQ);    // LDS column stride of the products
    static constexpr int IXP = ROWMAJ ? (IX + 63) / 64 * 64 : IX;  // product slots per quad row
    static constexpr int QC = TZ / 4;                          // new-level quads per column
    static constexpr int NQ1 = (WX * WQ + kThreads - 1) / kThreads;  // window quads per thread
    static constexpr int NQ2 = (IXP * PQ + kThreads - 1) / kThreads; // product quads per thread
    static constexpr int NQ3 = (BX * QC + kThreads - 1) / kThreads;  // new-level quads per thread
    static_assert(ROWMAJ || (TX * QC) % kThreads == 0, "box shape must give every thread whole quads");
};

struct TileInfo {
    int tile, sx, ex, sz, ez, wsx, wsz, plx, phx, plz, phz, pz0;
};

// new-level quad i of a box: column-major, so that a wavefront covers whole columns
// (contiguous 16-byte quads along z); row-major in thin z bands, where the quad row is
// uniform over a wavefront and is made a scalar
template <class S>
__device__ __forceinline__ void out_quad(const TileInfo& t, int i, int& gx, int& z0)
{
    if constexpr (S::ROWMAJ) {
        const int row = __builtin_amdgcn_readfirstlane(i / S::BX);
        gx = t.sx + (i - row * S::BX);
        z0 = t.sz + 4 * row;
    } else {
        gx = t.sx + i / S::QC;
        z0 = t.sz + 4 * (i % S::QC);
    }
}

// product quad of slot i: column lc of the IX product columns, quad row lq of PQ;
// false for a slot that holds none
template <class S>
__device__ __forceinline__ bool prod_slot(int i, int& lc, int& lq)
{
    if constexpr (S::ROWMAJ) {
        lq = __builtin_amdgcn_readfirstlane(i / S::IXP);
        lc = i - lq * S::IXP;
        return lc < S::IX && lq < S::PQ;
    } else {
        lc = i / S::PQ;
        lq = i - lc * S::PQ;
        return i < S::IX * S::PQ;
    }
}

struct StepSmem {
    float *ux, *uz, *mu, *mux, *px, *pz, *qx, *qz;
};

// D1 of a 12-cell column (cells z0-4 .. z0+7) at the 4 quad cells
__device__ __forceinline__ void d1_quad_z(const float* c, float* out, float c1, float c2)
{
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = c1 * (c[5 + k] - c[3 + k]) + c2 * (c[6 + k] - c[2 + k]);
}

// D2(c) of a 12-cell column at the 4 quad cells (2 fields sharing the coefficient)
__device__ __forceinline__ void d2_quad_z(const float* m, const float* u1, const float* u2, float* o1, float* o2)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const W5 w = d2_weights(&m[4 + k], 1);
        o1[k] = d2_apply(w, &u1[4 + k], 1);
        o2[k] = d2_apply(w, &u2[4 + k], 1);
    }
}

// i is a multiple of 4: the reads are indexed in quads, so that the compiler knows their
// 16-byte alignment and emits 128-bit LDS reads (not pairs of 64-bit ones, which conflict
// on banks for lanes 16 bytes apart)
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load_col12(const float* base, int i, float* c)
{
    const v4f* q = reinterpret_cast<const v4f*>(base) + (i >> 2);
    v4f a = q[-1], b = q[0], d = q[1];
    // whole quads: the compiler would otherwise read only the 8 rows a first derivative
    // uses, as 64-bit LDS reads 8 bytes off the 16-byte grid (2-way bank conflicts)
    __asm__ volatile("" : "+v"(a), "+v"(d));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        c[k] = a[k];
        c[4 + k] = b[k];
        c[8 + k] = d[k];
    }
}

// D2(c) u at global index g of n along stride s: interior weights or closure rows
__device__ __forceinline__ void d2_any2(const float* pu, const float* pv, const float* pc, int s, int g, int n,
                                        float& ou, float& ov)
{
    if (g >= 6 && g < n - 6) {
        const W5 w = d2_weights(pc, s);
        ou = d2_apply(w, pu, s);
        ov = d2_apply(w, pv, s);
    } else {
        d2_closure2(pu, pv, pc, s, g, n, ou, ov);
    }
}

__device__ __forceinline__ float d2_any(const float* pu, const float* pc, int s, int g, int n)
{
    if (g >= 6 && g < n - 6) return d2_apply(d2_weights(pc, s), pu, s);
    return d2_closure(pu, pc, s, g, n);
}

// ---- products of first derivatives on one quad --------------------------------------
template <bool VTI, class S, bool EDGE>
__device__ __forceinline__ void product_quad(const StepArgs& a, const StepSmem& sm, const TileInfo& t, int gx, int z0,
                                             float4 lam4, float4 lamx4, float4 c134, float* p, float* pz, float* qx,
                                             float* qz)
{
    const int nx = a.nx, nz = a.nz;
    const int w = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
    float dxux[4], dxuz[4], dzux[4], dzuz[4];
    if (!EDGE || (gx >= 4 && gx < nx - 4)) {
        float X[4][4], Z[4][4];
        const int off[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            unpack(ld4(&sm.ux[w + off[c] * S::WZ]), X[c]);
            unpack(ld4(&sm.uz[w + off[c] * S::WZ]), Z[c]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dxux[k] = a.c1x * (X[2][k] - X[1][k]) + a.c2x * (X[3][k] - X[0][k]);
            dxuz[k] = a.c1x * (Z[2][k] - Z[1][k]) + a.c2x * (Z[3][k] - Z[0][k]);
        }
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            dxux[k] = d1_closure(&sm.ux[w + k], S::WZ, gx, nx) * a.rdx;
            dxuz[k] = d1_closure(&sm.uz[w + k], S::WZ, gx, nx) * a.rdx;
        }
    }
    if (!EDGE || (z0 >= 4 && z0 + 3 < nz - 4)) {
        float Uc[12], Zc[12];
        load_col12(sm.ux, w, Uc);
        load_col12(sm.uz, w, Zc);
        d1_quad_z(Uc, dzux, a.c1z, a.c2z);
        d1_quad_z(Zc, dzuz, a.c1z, a.c2z);
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int gz = z0 + k;
            const bool in = gz >= 0 && gz < nz;
            dzux[k] = in ? d1<false>(&sm.ux[w + k], 1, gz, nz) * a.rdz : 0.f;
            dzuz[k] = in ? d1<false>(&sm.uz[w + k], 1, gz, nz) * a.rdz : 0.f;
        }
    }
    float mu[4], lam[4], lamx[4], c13[4];
    unpack(ld4(&sm.mu[w]), mu);
    unpack(lam4, lam);
    unpack(lamx4, lamx);
    unpack(c134, c13);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!VTI) {
            p[k] = lam[k] * (dxux[k] + dzuz[k]);
            pz[k] = 0.f;
        } else {
            p[k] = lamx[k] * dxux[k] + c13[k] * dzuz[k];
            pz[k] = lam[k] * dzuz[k] + c13[k] * dxux[k];
        }
        qx[k] = mu[k] * dxuz[k];
        qz[k] = mu[k] * dzux[k];
    }
}

// ---- SAT terms of one perimeter cell (we_op.cpp:726-822, 1705-1800) -------------------
template <bool VTI, class S>
__device__ __forceinline__ void sat_cell(const StepArgs& a, const float* sux, const float* suz, float mu0, int gx,
                                         int gz, float prx, float prz, float& rx, float& rz)
{
    const int nx = a.nx, nz = a.nz;
    const int g = gx * a.ldz + gz;
    const float lam0 = a.lam[g];
    const float rho0 = a.rho[g];
    const float f1tb = VTI ? a.c13[g] : lam0;
    const float epsf = VTI ? 1.f + 2.f * a.eps[g] : 1.f;
    for (int side = 0; side < 2; side++) {
        const int bc = side ? a.bc_bottom : a.bc_top;
        if (bc == 0 || gz != (side ? nz - 1 : 0)) continue;
        const float sgn = side ? 1.f : -1.f;
        const int sn = side ? -1 : 1;  // inward step along z
        float tt = sgn * mu0 * d1<false>(suz, S::WZ, gx, nx) * a.rdx + mu0 * bnd_deriv(sux, sn) * a.rdz;
        if (bc == 2) tt -= (0.5f * sqrtf(mu0 * rho0)) * prx / a.dt;
        rx -= a.sat_z * tt;
        tt = sgn * f1tb * d1<false>(sux, S::WZ, gx, nx) * a.rdx + (2.f * mu0) * bnd_deriv(suz, sn) * a.rdz;
        if (bc == 2) tt -= (0.5f * sqrtf((lam0 + 2.f * mu0) * rho0)) * prz / a.dt;
        rz -= a.sat_z * tt;
        rz += (a.sat_z * a.rdz) * lam0 * d1_row0(suz, sn);
    }
    for (int side = 0; side < 2; side++) {
        const int bc = side ? a.bc_right : a.bc_left;
        if (bc == 0 || gx != (side ? nx - 1 : 0)) continue;
        const float sgn = side ? 1.f : -1.f;
        const int sn = side ? -S::WZ : S::WZ;  // inward step along x
        const float f2 = epsf * (2.f * mu0);
        float tt = sgn * f1tb * d1<false>(suz, 1, gz, nz) * a.rdz + f2 * bnd_deriv(sux, sn) * a.rdx;
        if (bc == 2) tt -= (0.5f * sqrtf(epsf * (lam0 + 2.f * mu0) * rho0)) * prx / a.dt;
        rx -= a.sat_x * tt;
        rx += (a.sat_x * a.rdx) * (epsf * lam0) * d1_row0(sux, sn);
        tt = sgn * mu0 * d1<false>(sux, 1, gz, nz) * a.rdz + mu0 * bnd_deriv(suz, sn) * a.rdx;
        if (bc == 2) tt -= (0.5f * sqrtf(mu0 * rho0)) * prz / a.dt;
        rz -= a.sat_x * tt;
    }
}

// ---- right-hand side of one quad before sources: volume terms (we_op.cpp:690-720,
// 1665-1700) and SAT terms ----------------------------------------------------------
template <bool VTI, class S, bool EDGE, bool CW>
__device__ __forceinline__ void rhs_quad(const StepArgs& a, const StepSmem& sm, const TileInfo& t, int gx, int z0,
                                         const float* prx, const float* prz, float* rx, float* rz, const QuadW& qw)
{
    const int nx = a.nx, nz = a.nz;
    const int w = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
    const int q = (gx - t.plx) * S::PZ + (z0 - t.pz0);
    // x part: D2x[2mu or 2(1+2eps)mu] ux, D2x[mu] uz, Dx of the products
    if (!EDGE || (gx >= 6 && gx < nx - 6)) {
        float U[5][4], Z[5][4], M[CW ? 1 : 5][4], MX[VTI ? 5 : 1][4];
#pragma unroll
        for (int c = 0; c < 5; c++) {
            unpack(ld4(&sm.ux[w + (c - 2) * S::WZ]), U[c]);
            unpack(ld4(&sm.uz[w + (c - 2) * S::WZ]), Z[c]);
            if (!CW) unpack(ld4(&sm.mu[w + (c - 2) * S::WZ]), M[c]);
            if (VTI) unpack(ld4(&sm.mux[w + (c - 2) * S::WZ]), MX[c]);
        }
        float P[4][4], Q[4][4];
        const int off[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            unpack(ld4(&sm.px[q + off[c] * S::PZ]), P[c]);
            unpack(ld4(&sm.qz[q + off[c] * S::PZ]), Q[c]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float uk[5] = {U[0][k], U[1][k], U[2][k], U[3][k], U[4][k]};
            const float zk[5] = {Z[0][k], Z[1][k], Z[2][k], Z[3][k], Z[4][k]};
            W5 wx;
            if constexpr (CW) {
                wx = qw.x[k];
            } else {
                const float mk[5] = {M[CW ? 0 : 0][k], M[CW ? 0 : 1][k], M[CW ? 0 : 2][k], M[CW ? 0 : 3][k],
                                     M[CW ? 0 : 4][k]};
                wx = d2_weights(&mk[2], 1);
            }
            float d2u;
            if (VTI) {
                const float xk[5] = {MX[0][k], MX[1][k], MX[2][k], MX[3][k], MX[4][k]};
                d2u = d2_apply(d2_weights(&xk[2], 1), &uk[2], 1);
            } else {
                d2u = d2_apply(wx, &uk[2], 1);
            }
            if constexpr (CW) {  // cached weights already carry 1/dx^2
                rx[k] = 2.f * d2u;
                rz[k] = d2_apply(wx, &zk[2], 1);
            } else {
                rx[k] = 2.f * d2u * a.rdx2;
                rz[k] = d2_apply(wx, &zk[2], 1) * a.rdx2;
            }
            rx[k] += a.c1x * (P[2][k] - P[1][k]) + a.c2x * (P[3][k] - P[0][k]);
            rz[k] += a.c1x * (Q[2][k] - Q[1][k]) + a.c2x * (Q[3][k] - Q[0][k]);
        }
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            if (VTI) {
                rx[k] = 2.f * d2_closure(&sm.ux[w + k], &sm.mux[w + k], S::WZ, gx, nx) * a.rdx2;
                rz[k] = d2_closure(&sm.uz[w + k], &sm.mu[w + k], S::WZ, gx, nx) * a.rdx2;
            } else {
                float dxx, dxz;
                d2_closure2(&sm.ux[w + k], &sm.uz[w + k], &sm.mu[w + k], S::WZ, gx, nx, dxx, dxz);
                rx[k] = 2.f * dxx * a.rdx2;
                rz[k] = dxz * a.rdx2;
            }
            rx[k] += d1<false>(&sm.px[q + k], S::PZ, gx, nx) * a.rdx;
            rz[k] += d1<false>(&sm.qz[q + k], S::PZ, gx, nx) * a.rdx;
        }
    }
    // z part: D2z[2mu] uz, D2z[mu] ux, Dz of the products
    if (!EDGE || (z0 >= 6 && z0 + 3 < nz - 6)) {
        float Uc[12], Zc[12], d2zx[4], d2zz[4];
        load_col12(sm.ux, w, Uc);
        load_col12(sm.uz, w, Zc);
        if constexpr (CW) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                d2zz[k] = d2_apply(qw.z[k], &Zc[4 + k], 1);
                d2zx[k] = d2_apply(qw.z[k], &Uc[4 + k], 1);
            }
        } else {
            float Mc[12];
            load_col12(sm.mu, w, Mc);
            d2_quad_z(Mc, Zc, Uc, d2zz, d2zx);
        }
        float Pc[12], Qc[12], dpz[4], dqz[4];
        load_col12(VTI ? sm.pz : sm.px, q, Pc);
        load_col12(sm.qx, q, Qc);
        d1_quad_z(Pc, dpz, a.c1z, a.c2z);
        d1_quad_z(Qc, dqz, a.c1z, a.c2z);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if constexpr (CW) {  // cached weights already carry 1/dz^2
                rz[k] += 2.f * d2zz[k];
                rx[k] += d2zx[k];
            } else {
                rz[k] += 2.f * d2zz[k] * a.rdz2;
                rx[k] += d2zx[k] * a.rdz2;
            }
            rz[k] += dpz[k];
            rx[k] += dqz[k];
        }
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int gz = z0 + k;
            if (gz >= nz) continue;
            float dzz, dzx;
            d2_any2(&sm.uz[w + k], &sm.ux[w + k], &sm.mu[w + k], 1, gz, nz, dzz, dzx);
            rz[k] += 2.f * dzz * a.rdz2;
            rx[k] += dzx * a.rdz2;
            rz[k] += d1<false>(VTI ? &sm.pz[q + k] : &sm.px[q + k], 1, gz, nz) * a.rdz;
            rx[k] += d1<false>(&sm.qx[q + k], 1, gz, nz) * a.rdz;
        }
    }
    // SAT lines
    if (EDGE && (gx == 0 || gx == nx - 1 || z0 == 0 || z0 + 3 >= nz - 1)) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int gz = z0 + k;
            if (gz >= nz) continue;
            if (gz == 0 || gz == nz - 1 || gx == 0 || gx == nx - 1)
                sat_cell<VTI, S>(a, &sm.ux[w + k], &sm.uz[w + k], sm.mu[w + k], gx, gz, prx[k], prz[k], rx[k], rz[k]);
        }
    }
}

// First derivatives of a pair of fields (u, v) at the 4 cells of a quad from an LDS
// array with column stride CS: (dx u, dz u, dx v, dz v), scaled by 1/h.
template <bool EDGE, int CS>
__device__ __forceinline__ void quad_derivs(const StepArgs& a, const float* u, const float* v, int w, int gx, int z0,
                                            float* xu, float* zu, float* xv, float* zv)
{
    const int nx = a.nx, nz = a.nz;
    const float rdx = a.rdx, rdz = a.rdz;
    if (!EDGE || (gx >= 4 && gx < nx - 4)) {
        float U[4][4], V[4][4];
        const int off[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            unpack(ld4(&u[w + off[c] * CS]), U[c]);
            unpack(ld4(&v[w + off[c] * CS]), V[c]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xu[k] = a.c1x * (U[2][k] - U[1][k]) + a.c2x * (U[3][k] - U[0][k]);
            xv[k] = a.c1x * (V[2][k] - V[1][k]) + a.c2x * (V[3][k] - V[0][k]);
        }
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            xu[k] = d1_closure(&u[w + k], CS, gx, nx) * rdx;
            xv[k] = d1_closure(&v[w + k], CS, gx, nx) * rdx;
        }
    }
    if (!EDGE || (z0 >= 4 && z0 + 3 < nz - 4)) {
        float Uc[12], Vc[12];
        load_col12(u, w, Uc);
        load_col12(v, w, Vc);
        d1_quad_z(Uc, zu, a.c1z, a.c2z);
        d1_quad_z(Vc, zv, a.c1z, a.c2z);
    } else if constexpr (EDGE) {
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const int gz = z0 + k;
            const bool in = gz < nz;
            zu[k] = in ? d1<false>(&u[w + k], 1, gz, nz) * rdz : 0.f;
            zv[k] = in ? d1<false>(&v[w + k], 1, gz, nz) * rdz : 0.f;
        }
    }
}

// Imaging condition of one quad (compute_gradients, we_op.cpp:519-561): a* are the
// derivatives of the lagged adjoint level (x_x, x_z, z_x, z_z), t* those of forward
// snapshot idx (x_x, z_z, z_x, x_z), f0/fm/fp the snapshots idx, idx-1, idx+1 at the
// cells, o the adjoint level of the density term.  Adds to the accumulators of the
// lambda, mu, rho gradients (VTI stiffness mode: c11, c13, c33, c55, rho).
// Density term of the imaging condition of one quad: o is the adjoint level after the
// step, f0/fm/fp the snapshots idx, idx-1, idx+1 at the cells (second time difference;
// one-sided at the first and last snapshot, mode 1 and 2).  It goes to gr, or to g4 in
// the VTI stiffness mode.  (The accumulators are updated unconditionally with selected
// values: a branch on the mode that picks the array would put the arrays in scratch.)
template <bool VTI>
__device__ __forceinline__ void image_rho(const StepArgs& a, int z0, const float* f0x, const float* f0z,
                                          const float* fmx, const float* fmz, const float* fpx, const float* fpz,
                                          const float* ox, const float* oz, float* gr, float* g4)
{
    const bool st = VTI && a.stiff;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (z0 + k >= a.nz) continue;
        float d2x, d2z;
        if (a.img_mode == 0) {
            d2x = fpx[k] - 2.f * f0x[k] + fmx[k];
            d2z = fpz[k] - 2.f * f0z[k] + fmz[k];
        } else if (a.img_mode == 1) {
            d2x = fpx[k] - f0x[k];
            d2z = fpz[k] - f0z[k];
        } else {
            d2x = -f0x[k] + fmx[k];
            d2z = -f0z[k] + fmz[k];
        }
        const float r = a.wrho * (ox[k] * d2x + oz[k] * d2z);
        if (VTI) {
            gr[k] += st ? 0.f : r;
            g4[k] += st ? r : 0.f;
        } else {
            gr[k] += r;
        }
    }
}

// Stiffness terms of the imaging condition of one quad: a* are the derivatives of the
// lagged adjoint level (x_x, x_z, z_x, z_z), t* those of forward snapshot idx (x_x,
// z_z, z_x, x_z).  Adds to the lambda and mu accumulators gl, gm; in the VTI stiffness
// mode to the c11, c13, c33, c55 accumulators gl, gm, gr, g3.
template <bool VTI>
__device__ __forceinline__ void image_stiff(const StepArgs& a, long long g, int z0, const float* axx, const float* azx,
                                            const float* axz, const float* azz, const float* t0x, const float* t1z,
                                            const float* t2x, const float* t3z, float* gl, float* gm, float* gr,
                                            float* g3)
{
    const bool st = VTI && a.stiff;
    float e[4], v2[4], v3[4];
    if (VTI && !st) {
        unpack(ld4(a.ge + g), e);
        unpack(ld4(a.gv2 + g), v2);
        unpack(ld4(a.gv3 + g), v3);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (z0 + k >= a.nz) continue;
        const float shear = (azx[k] + axz[k]) * (t2x[k] + t3z[k]);
        if (!VTI) {
            gl[k] += a.wlm * (axx[k] + azz[k]) * (t0x[k] + t1z[k]);
            gm[k] += a.wlm * (shear + 2.f * axx[k] * t0x[k] + 2.f * azz[k] * t1z[k]);
        } else {
            const float xx = axx[k] * t0x[k], zz = azz[k] * t1z[k];
            const float cr = axx[k] * t1z[k] + azz[k] * t0x[k];
            // stiffness mode: d/dc11, d/dc13, d/dc33, d/dc55 of the same bilinear form
            const float dl = st ? xx : e[k] * xx + zz + v2[k] * cr;
            const float dm = st ? cr : shear + 2.f * e[k] * xx + 2.f * zz + v3[k] * cr;
            gl[k] += a.wlm * dl;
            gm[k] += a.wlm * dm;
            gr[k] += st ? a.wlm * zz : 0.f;
            g3[k] += st ? a.wlm * shear : 0.f;
        }
    }
}

// Imaging condition of one quad (compute_gradients, we_op.cpp:519-561) into the
// accumulators of the lambda, mu, rho gradients (VTI stiffness mode: c11, c13, c33, c55,
// rho in gl, gm, gr, g3, g4).
template <bool VTI>
__device__ __forceinline__ void image_accumulate(const StepArgs& a, long long g, int z0, const float* axx,
                                                 const float* azx, const float* axz, const float* azz,
                                                 const float* t0x, const float* t1z, const float* t2x,
                                                 const float* t3z, const float* f0x, const float* f0z,
                                                 const float* fmx, const float* fmz, const float* fpx,
                                                 const float* fpz, const float* ox, const float* oz, float* gl,
                                                 float* gm, float* gr, float* g3, float* g4)
{
    image_stiff<VTI>(a, g, z0, axx, azx, axz, azz, t0x, t1z, t2x, t3z, gl, gm, gr, g3);
    image_rho<VTI>(a, z0, f0x, f0z, fmx, fmz, fpx, fpz, ox, oz, gr, g4);
}

// Hides a value from the optimiser so that address arithmetic derived from it is
// recomputed inside the shot loop instead of being hoisted out and kept live in
// registers across it (which multiplies the register footprint of the loop body).
__device__ __forceinline__ void opaque(int& v) { __asm__ volatile("" : "+s"(v)); }
__device__ __forceinline__ void opaque_v(int& v) { __asm__ volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque(TileInfo& t)
{
    opaque(t.tile);
    opaque(t.sx);
    opaque(t.ex);
    opaque(t.sz);
    opaque(t.ez);
    opaque(t.wsx);
    opaque(t.wsz);
    opaque(t.plx);
    opaque(t.phx);
    opaque(t.plz);
    opaque(t.phz);
    opaque(t.pz0);
}

// 16 zero bytes: the source of direct-to-LDS window loads at positions outside the grid
__device__ __attribute__((aligned(16))) float kZeroQuad[4] = {0.f, 0.f, 0.f, 0.f};

// One quad of a window straight from global memory into LDS (global_load_lds_dwordx4):
// lane l of the wavefront writes the 16 bytes at lds_wave + 16 l, lds_wave being the
// same for all lanes of the wavefront.
__device__ __forceinline__ void quad_to_lds(const float* src, float* lds_wave)
{
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_wave, 16, 0, 0);
}

// A workgroup advances the shots of one group through one box of the grid.  Boxes of the
// core (EDGE = false) hold only cells whose stencils are all interior, so no closure or
// SAT code is compiled in; boxes of the edge bands (EDGE = true) take the closure and
// SAT paths where a quad needs them.
// IMG (the image_fused option; by default the imaging runs in k_img): adjoint step it with
// the imaging by parts in time fused in (see k_img with SBP): the density term of
// snapshot j = f0 against wn next + wc curr + wp prev at the cells (prev as stored, before
// the step's taper; next before the step's point sources, whose share k_inject_cells
// adds), and the stiffness terms of snapshot j-1 = fm, whose derivatives meet those of the
// step's current level (the reference's lag).  fm's window goes through the LDS of the
// products once the new level is done.  The gradient contributions of the workgroup's
// shots are summed in registers and added once per launch to the gradient slab of the
// group's first shot.
template <bool VTI, int TX, int TZ, bool EDGE, bool IMG>
__global__ void __launch_bounds__(kThreads)
k_step(const StepArgs a, const int4* boxes, const int* list)
{
    using S = StepShape<TX, TZ>;
    __shared__ __attribute__((aligned(16))) float sUx[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sUz[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sMu[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sMuX[VTI ? S::WX * S::WZ : 4];  // (1+2eps) mu
    __shared__ __attribute__((aligned(16))) float sPx[S::IX * S::PZ];
    __shared__ __attribute__((aligned(16))) float sPz[VTI ? S::IX * S::PZ : 4];
    __shared__ __attribute__((aligned(16))) float sQx[S::IX * S::PZ];
    __shared__ __attribute__((aligned(16))) float sQz[S::IX * S::PZ];
    StepSmem sm{sUx, sUz, sMu, sMuX, sPx, sPz, sQx, sQz};
    // IMG: snapshot fm's window in the product layout, over the products (dead by then)
    float* const sFx = sPx;
    float* const sFz = sQx;

    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    TileInfo t;
    const int slot = lin / a.ngroups;
    const int group = lin - slot * a.ngroups;
    t.tile = list[slot];
    const int4 box = boxes[t.tile];
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    t.sx = box.x;
    t.ex = box.y;
    t.sz = box.z;
    t.ez = box.w;
    t.wsx = window_start(t.sx, S::HX);
    if constexpr (S::COMPACT) {
        t.wsz = t.sz == 0 ? 0 : t.sz - 8;
        t.pz0 = t.sz == 0 ? 0 : t.sz - 4;
    } else {
        t.wsz = t.sz - S::HZ;
        t.pz0 = t.sz - 4;
    }
    product_range(t.sx, t.ex, nx, t.plx, t.phx);
    product_range(t.sz, t.ez, nz, t.plz, t.phz);
    const long long nxz = (long long)nx * ldz;
    const int tid = threadIdx.x;
    const int s0 = group * a.spb;
    const int s1 = min(a.nshots, s0 + a.spb);

    // ---- per tile: coefficient window, product coefficients, new-level factors ----
#pragma unroll
    for (int m = 0; m < S::NQ1; m++) {
        const int i = tid + m * kThreads;
        if (i < S::WX * S::WQ) {
            const int lx = i / S::WQ, lq = i - lx * S::WQ;
            const int gx = t.wsx + lx, gz = t.wsz + 4 * lq;
            const bool ok = gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
            const int g = ok ? gx * ldz + gz : 0;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            const int o = lx * S::WZ + 4 * lq;
            st4(&sMu[o], ok ? ld4(a.mu + g) : z4);
            if (VTI) st4(&sMuX[o], ok ? ld4(a.mux + g) : z4);
        }
    }
    float4 plam[S::NQ2], plamx[VTI ? S::NQ2 : 1], pc13[VTI ? S::NQ2 : 1];
#pragma unroll
    for (int m = 0; m < S::NQ2; m++) {
        const int i = tid + m * kThreads;
        int lc, lq;
        const bool slot = prod_slot<S>(i, lc, lq);
        const int gx = t.plx + lc, z0 = t.pz0 + 4 * lq;
        const bool ok = slot && gx < t.phx && z0 + 3 >= t.plz && z0 < t.phz;
        const int g = ok ? gx * ldz + z0 : 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        plam[m] = ok ? ld4(a.lam + g) : z4;
        if (VTI) {
            plamx[m] = ok ? ld4(a.lamx + g) : z4;
            pc13[m] = ok ? ld4(a.c13 + g) : z4;
        }
    }
    float4 qir[S::NQ3], qtz[S::NQ3];
    float qtx[S::NQ3];
#pragma unroll
    for (int m = 0; m < S::NQ3; m++) {
        const int i = tid + m * kThreads;
        int gx, gz;
        out_quad<S>(t, i, gx, gz);
        const bool ok = gx < t.ex && gz < t.ez;
        const int g = ok ? gx * ldz + gz : 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        qir[m] = ok ? ld4(a.irho + g) : z4;
        qtz[m] = ok ? ld4(a.taper_z + gz) : z4;
        qtx[m] = ok ? a.taper_x[gx] : 0.f;
    }

    // ---- D2 weights of the new-level quads (core boxes, isotropic): once per workgroup ----
    // (not with IMG: the imaging's accumulators and second phase need the registers)
    constexpr bool CW = !EDGE && !VTI && S::NQ3 == 1 && !IMG;
    QuadW qw[CW ? S::NQ3 : 1];
    if constexpr (CW) {
        __syncthreads();  // the coefficient window is in LDS
#pragma unroll
        for (int m = 0; m < S::NQ3; m++) {
            const int i = tid + m * kThreads;
            int gx, z0;
            out_quad<S>(t, i, gx, z0);
            if (gx >= t.ex || z0 >= t.ez) continue;
            const int w = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
            float M[5][4], Mc[12];
#pragma unroll
            for (int c = 0; c < 5; c++) unpack(ld4(&sMu[w + (c - 2) * S::WZ]), M[c]);
            load_col12(sMu, w, Mc);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float mk[5] = {M[0][k], M[1][k], M[2][k], M[3][k], M[4][k]};
                qw[m].x[k] = d2_scaled(d2_weights(&mk[2], 1), a.rdx2);
                qw[m].z[k] = d2_scaled(d2_weights(&Mc[4 + k], 1), a.rdz2);
            }
        }
    }

    // ---- state window of one shot into registers ----
    float4 wux[S::NQ1], wuz[S::NQ1];
    auto load_window = [&](int shot, int tid, const TileInfo& t) {
        const float* cux = a.curr + shot * a.shot_stride;
#pragma unroll
        for (int m = 0; m < S::NQ1; m++) {
            const int i = tid + m * kThreads;
            const int lx = i / S::WQ, lq = i - lx * S::WQ;
            const int gx = t.wsx + lx, gz = t.wsz + 4 * lq;
            const bool ok = i < S::WX * S::WQ && gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
            const int g = ok ? gx * ldz + gz : 0;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            wux[m] = ok ? ld4(cux + g) : z4;
            wuz[m] = ok ? ld4(cux + nxz + g) : z4;
        }
    };
    // core boxes: the state window goes straight from global memory into LDS at the start
    // of each shot (no register staging); other boxes prefetch the next shot's window
    // into registers during this shot
    constexpr bool DMA = !EDGE && S::WZ == 4 * S::WQ;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (!DMA && s0 < s1) load_window(s0, tid, t);

    float gl[IMG ? S::NQ3 : 1][4], gm[IMG ? S::NQ3 : 1][4], gr[IMG ? S::NQ3 : 1][4];
    float g3[IMG && VTI ? S::NQ3 : 1][4], g4[IMG && VTI ? S::NQ3 : 1][4];  // VTI stiffness mode
    if constexpr (IMG) {
#pragma unroll
        for (int m = 0; m < S::NQ3; m++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                gl[m][k] = gm[m][k] = gr[m][k] = 0.f;
                if (VTI) g3[m][k] = g4[m][k] = 0.f;
            }
    }

    const TileInfo t0 = t;
    const int tid0 = tid;
    for (int shot = s0; shot < s1; shot++) {
        TileInfo t = t0;
        int tid = tid0;
        opaque(t);
        opaque_v(tid);
        // previous level of this shot's new-level quads
        float4 qprx[S::NQ3], qprz[S::NQ3];
        const float* pux = a.prev + shot * a.shot_stride;
#pragma unroll
        for (int m = 0; m < S::NQ3; m++) {
            const int i = tid + m * kThreads;
            int gx, gz;
        out_quad<S>(t, i, gx, gz);
            const bool ok = gx < t.ex && gz < t.ez;
            const int g = ok ? gx * ldz + gz : 0;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            qprx[m] = ok ? ld4(pux + g) : z4;
            qprz[m] = ok ? ld4(pux + nxz + g) : z4;
        }
        __syncthreads();  // the previous shot is done with the LDS state and products
        if constexpr (DMA) {
            const float* cux = a.curr + shot * a.shot_stride;
#pragma unroll
            for (int m = 0; m < S::NQ1; m++) {
                const int i = tid + m * kThreads;
                if (i < S::WX * S::WQ) {
                    const int lx = i / S::WQ, lq = i - lx * S::WQ;
                    const int gx = t.wsx + lx, gz = t.wsz + 4 * lq;
                    const bool ok = gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
                    const int g = gx * ldz + gz;
                    const int o = 4 * (m * kThreads + wave * 64);  // window quad i at 4 i
                    quad_to_lds(ok ? cux + g : kZeroQuad, &sUx[o]);
                    quad_to_lds(ok ? cux + nxz + g : kZeroQuad, &sUz[o]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        } else {
#pragma unroll
            for (int m = 0; m < S::NQ1; m++) {
                const int i = tid + m * kThreads;
                if (i < S::WX * S::WQ) {
                    const int lx = i / S::WQ, lq = i - lx * S::WQ;
                    const int o = lx * S::WZ + 4 * lq;
                    st4(&sUx[o], wux[m]);
                    st4(&sUz[o], wuz[m]);
                }
            }
        }
        __syncthreads();
        if (!DMA && shot + 1 < s1) load_window(shot + 1, tid, t);

        // products of first derivatives
#pragma unroll
        for (int m = 0; m < S::NQ2; m++) {
            const int i = tid + m * kThreads;
            int lc, lq;
            if (!prod_slot<S>(i, lc, lq)) continue;
            const int gx = t.plx + lc, z0 = t.pz0 + 4 * lq;
            float p[4] = {0.f, 0.f, 0.f, 0.f}, pz[4] = {0.f, 0.f, 0.f, 0.f}, qx[4] = {0.f, 0.f, 0.f, 0.f},
                  qz[4] = {0.f, 0.f, 0.f, 0.f};
            if (gx < t.phx && z0 + 3 >= t.plz && z0 < t.phz)
                product_quad<VTI, S, EDGE>(a, sm, t, gx, z0, plam[m], VTI ? plamx[m] : plam[m], VTI ? pc13[m] : plam[m],
                                      p, pz, qx, qz);
            const int o = lc * S::PZ + 4 * lq;
            st4(&sPx[o], make_float4(p[0], p[1], p[2], p[3]));
            if (VTI) st4(&sPz[o], make_float4(pz[0], pz[1], pz[2], pz[3]));
            st4(&sQx[o], make_float4(qx[0], qx[1], qx[2], qx[3]));
            st4(&sQz[o], make_float4(qz[0], qz[1], qz[2], qz[3]));
        }
        __syncthreads();

        // new level
        float* nux = a.next + shot * a.shot_stride;
        float* nuz = nux + nxz;
#pragma unroll
        for (int m = 0; m < S::NQ3; m++) {
            const int i = tid + m * kThreads;
            int gx, z0;
            out_quad<S>(t, i, gx, z0);
            if (gx >= t.ex || z0 >= t.ez) continue;
            float prx[4], prz[4], ir[4], tz[4];
            unpack(qprx[m], prx);
            unpack(qprz[m], prz);
            unpack(qir[m], ir);
            unpack(qtz[m], tz);
            const float txf = qtx[m];
            if (a.taper_prev) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    prx[k] = (prx[k] * tz[k]) * txf;
                    prz[k] = (prz[k] * tz[k]) * txf;
                }
            }
            float rx[4], rz[4];
            rhs_quad<VTI, S, EDGE, CW>(a, sm, t, gx, z0, prx, prz, rx, rz, qw[CW ? m : 0]);
            const int wq = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
            float cx[4], cz[4], vx[4], vz[4];
            unpack(ld4(&sUx[wq]), cx);
            unpack(ld4(&sUz[wq]), cz);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // leapfrog (we_op.cpp:824-830) with ir = dt^2 / rho
                vx[k] = ir[k] * rx[k] + (2.f * cx[k] - prx[k]);
                vz[k] = ir[k] * rz[k] + (2.f * cz[k] - prz[k]);
            }
            if (EDGE && (gx == 0 || gx == nx - 1 || z0 == 0 || z0 + 3 >= nz - 1)) {
                // implicit part of the absorbing SAT on the perimeter (spatial_operators.cpp:307-521)
                for (int k = 0; k < 4; k++) {
                    const int gz = z0 + k;
                    if (gz >= nz) continue;
                    const bool edge = (gz == 0) | (gz == nz - 1) | (gx == 0) | (gx == nx - 1);
                    if (!edge) continue;
                    float dvx, dvz;
                    if (gz == 0) { dvx = a.div_tb[gx]; dvz = a.div_tb[nx + gx]; }
                    else if (gz == nz - 1) { dvx = a.div_tb[2 * nx + gx]; dvz = a.div_tb[3 * nx + gx]; }
                    else if (gx == 0) { dvx = a.div_lr[gz]; dvz = a.div_lr[nz + gz]; }
                    else { dvx = a.div_lr[2 * nz + gz]; dvz = a.div_lr[3 * nz + gz]; }
                    vx[k] = vx[k] / dvx;
                    vz[k] = vz[k] / dvz;
                }
            }
            // cos^2 taper of the new level (we_op.cpp:846-861); cells past nz stay zero
            float ox[4], oz[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool in = z0 + k < nz;
                ox[k] = in ? (vx[k] * tz[k]) * txf : 0.f;
                oz[k] = in ? (vz[k] * tz[k]) * txf : 0.f;
            }
            const int g = gx * ldz + z0;
            st4(nux + g, make_float4(ox[0], ox[1], ox[2], ox[3]));
            st4(nuz + g, make_float4(oz[0], oz[1], oz[2], oz[3]));
            if constexpr (IMG) {
                // density term of snapshot j: f0 against the weighted adjoint levels
                __builtin_amdgcn_sched_barrier(0);
                const long long fo = shot * a.fshot_stride + g;
                float f0x[4], f0z[4], pvx[4], pvz[4];
                unpack(ld4(a.f0 + fo), f0x);
                unpack(ld4(a.f0 + fo + nxz), f0z);
                unpack(qprx[m], pvx);
                unpack(qprz[m], pvz);
                const bool st = VTI && a.stiff;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float ax = a.wn * ox[k] + a.wc * cx[k] + a.wp * pvx[k];
                    const float az = a.wn * oz[k] + a.wc * cz[k] + a.wp * pvz[k];
                    const float r = z0 + k < nz ? f0x[k] * ax + f0z[k] * az : 0.f;
                    if (VTI) {
                        gr[m][k] += st ? 0.f : r;
                        g4[VTI ? m : 0][k] += st ? r : 0.f;
                    } else {
                        gr[m][k] += r;
                    }
                }
            }
        }
        if constexpr (IMG) {
            // stiffness terms of snapshot j-1: its window replaces the products
            __syncthreads();  // every wavefront is done with the products
            const float* Fm = a.fm + shot * a.fshot_stride;
#pragma unroll
            for (int m = 0; m < S::NQ2; m++) {
                const int i = tid + m * kThreads;
                int lc, lq;
                const bool slot = prod_slot<S>(i, lc, lq);
                const int gx = t.plx + lc, z0 = t.pz0 + 4 * lq;
                const bool ok = slot && gx < nx && z0 >= 0 && z0 < ldz;
                const int g = ok ? gx * ldz + z0 : 0;
                if constexpr (DMA && !S::ROWMAJ) {
                    // quad i at float 4 i (column-major product slots), as the window DMA
                    if (slot) {
                        const int o = 4 * (m * kThreads + wave * 64);
                        quad_to_lds(ok ? Fm + g : kZeroQuad, &sFx[o]);
                        quad_to_lds(ok ? Fm + nxz + g : kZeroQuad, &sFz[o]);
                    }
                } else {
                    if (!slot) continue;
                    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    const int o = lc * S::PZ + 4 * lq;
                    st4(&sFx[o], ok ? ld4(Fm + g) : z4);
                    st4(&sFz[o], ok ? ld4(Fm + nxz + g) : z4);
                }
            }
            if constexpr (DMA && !S::ROWMAJ) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
#pragma unroll
            for (int m = 0; m < S::NQ3; m++) {
                const int i = tid + m * kThreads;
                int gx, z0;
                out_quad<S>(t, i, gx, z0);
                if (gx >= t.ex || z0 >= t.ez) continue;
                const int wq = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
                const int q = (gx - t.plx) * S::PZ + (z0 - t.pz0);
                // adjoint derivatives of the step's current level, forward derivatives of
                // snapshot j-1 (we_op.cpp:519-541)
                float axx[4], azx[4], axz[4], azz[4], t0x[4], t1z[4], t2x[4], t3z[4];
                quad_derivs<EDGE, S::WZ>(a, sUx, sUz, wq, gx, z0, axx, azx, axz, azz);
                quad_derivs<EDGE, S::PZ>(a, sFx, sFz, q, gx, z0, t0x, t3z, t2x, t1z);
                image_stiff<VTI>(a, (long long)gx * ldz + z0, z0, axx, azx, axz, azz, t0x, t1z, t2x, t3z, gl[m],
                                 gm[m], gr[m], g3[VTI ? m : 0]);
            }
        }
    }
    if constexpr (IMG) {
        float* G = a.grad + s0 * a.gslab_stride;
#pragma unroll
        for (int m = 0; m < S::NQ3; m++) {
            const int i = tid + m * kThreads;
            int gx, z0;
            out_quad<S>(t, i, gx, z0);
            if (gx >= t.ex || z0 >= t.ez || s0 >= s1) continue;
            const long long g = (long long)gx * ldz + z0;
            float4 v = ld4(G + g);
            st4(G + g, make_float4(v.x + gl[m][0], v.y + gl[m][1], v.z + gl[m][2], v.w + gl[m][3]));
            v = ld4(G + nxz + g);
            st4(G + nxz + g, make_float4(v.x + gm[m][0], v.y + gm[m][1], v.z + gm[m][2], v.w + gm[m][3]));
            v = ld4(G + 2 * nxz + g);
            st4(G + 2 * nxz + g, make_float4(v.x + gr[m][0], v.y + gr[m][1], v.z + gr[m][2], v.w + gr[m][3]));
            if (VTI && a.stiff) {
                v = ld4(G + 3 * nxz + g);
                st4(G + 3 * nxz + g, make_float4(v.x + g3[m][0], v.y + g3[m][1], v.z + g3[m][2], v.w + g3[m][3]));
                v = ld4(G + 4 * nxz + g);
                st4(G + 4 * nxz + g, make_float4(v.x + g4[m][0], v.y + g4[m][1], v.z + g4[m][2], v.w + g4[m][3]));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Imaging condition after an adjoint step (compute_gradients, we_op.cpp:505-562).
// a.curr is the lagged adjoint level (its first derivatives are imaged), a.next the
// level after the step and its point sources (the density term), a.f0/fm/fp forward
// snapshots idx, idx-1, idx+1.  A workgroup owns a TX x 64 tile of the tile grid and a
// group of a.spb shots; per shot the adjoint and snapshot windows (halo 2 in x, 4 in z)
// go through LDS, the next shot's windows are loaded into registers meanwhile, and the
// contributions are summed in registers and added once to the gradient slab of the
// group's first shot.  Tiles with every cell at least 4 from the edges take the interior
// stencils only; the others choose per quad.
// ---------------------------------------------------------------------------
// SBP = true: the density term by parts in time (a.aprev is the derivative level, the
// cell values are the adjoint levels it and it+1, see StepArgs).
template <bool VTI, int TX, int NBUF, bool SBP>
__global__ void __launch_bounds__(kThreads) k_img(const StepArgs a)
{
    constexpr int TZ = kTileZ, HX = 2, HZ = 4;
    constexpr int WX = TX + 2 * HX, WZ = TZ + 2 * HZ, WQ = WZ / 4;
    constexpr int NQ1 = (WX * WQ + kThreads - 1) / kThreads;
    constexpr int QC = TZ / 4, NQ3 = TX * QC / kThreads;
    constexpr int WS = WX * WZ;  // floats per window array
    static_assert((TX * QC) % kThreads == 0, "tile must give every thread whole quads");
    // [buffer][Ax, Az, Fx, Fz][window]: window quad i of an array at float offset 4 i, so
    // that the quads a wavefront loads are consecutive in LDS
    __shared__ __attribute__((aligned(16))) float sW[NBUF * 4 * WS];

    // group-major order: the tiles of one shot group follow each other (z first), so that
    // the window halos a tile shares with its neighbours are still in the XCD's L2
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int ntiles = a.ntx * a.ntz;
    const int group = lin / ntiles;
    const int tile = lin - group * ntiles;
    const int tix = tile / a.ntz, tiz = tile - tix * a.ntz;
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    int sx, ex, sz, ez;
    tile_range_x_hd(tix, a.ntx, TX, nx, sx, ex);
    tile_range_z_hd(tiz, a.ntz, nz, ldz, sz, ez);
    const int wsx = window_start(sx, HX), wsz = sz - HZ;
    const bool inner = sx >= 4 && ex <= nx - 4 && sz >= 4 && ez <= nz - 4;
    const long long nxz = (long long)nx * ldz;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int s0 = group * a.spb;
    const int s1 = min(a.nshots, s0 + a.spb);

    // windows of one shot into buffer b: window quads tid + m * 256
    auto issue_windows = [&](int shot, int b) {
        const float* A = (SBP ? a.aprev : a.curr) + shot * a.shot_stride;
        const float* F = a.f0 + shot * a.fshot_stride;
        float* base = sW + b * 4 * WS;
#pragma unroll
        for (int m = 0; m < NQ1; m++) {
            const int i = tid + m * kThreads;
            if (i < WX * WQ) {
                const int lx = i / WQ, lq = i - lx * WQ;
                const int gx = wsx + lx, gz = wsz + 4 * lq;
                const bool ok = gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
                const long long g = (long long)gx * ldz + gz;
                float* dst = base + 4 * (m * kThreads + wave * 64);
                quad_to_lds(ok ? A + g : kZeroQuad, dst);
                quad_to_lds(ok ? A + nxz + g : kZeroQuad, dst + WS);
                quad_to_lds(ok ? F + g : kZeroQuad, dst + 2 * WS);
                quad_to_lds(ok ? F + nxz + g : kZeroQuad, dst + 3 * WS);
            }
        }
    };

    float gl[NQ3][4], gm[NQ3][4], gr[NQ3][4], g3[VTI ? NQ3 : 1][4], g4[VTI ? NQ3 : 1][4];
#pragma unroll
    for (int m = 0; m < NQ3; m++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            gl[m][k] = gm[m][k] = gr[m][k] = 0.f;
            if (VTI) g3[m][k] = g4[m][k] = 0.f;
        }

    if (NBUF == 2 && s0 < s1) issue_windows(s0, 0);
    for (int shot = s0; shot < s1; shot++) {
        const int b = NBUF == 2 ? (shot - s0) & 1 : 0;
        if (NBUF == 1) {
            if (shot > s0) __syncthreads();  // every wavefront is done with the windows
            issue_windows(shot, 0);
        }
        // cell values of this shot: the adjoint density level and snapshots idx-1, idx+1
        // (by parts: the adjoint levels it+1 and it)
        float4 qnx[NQ3], qnz[NQ3], qmx[NQ3], qmz[NQ3], qpx[NQ3], qpz[SBP ? 1 : NQ3];
#pragma unroll
        for (int m = 0; m < NQ3; m++) {
            const int i = tid + m * kThreads;
            const int gx = sx + i / QC, z0 = sz + 4 * (i % QC);
            const bool ok = gx < ex && z0 < ez;
            const long long g = ok ? (long long)gx * ldz + z0 : 0;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            const float* N = a.next + shot * a.shot_stride;
            qnx[m] = ok ? ld4(N + g) : z4;
            qnz[m] = ok ? ld4(N + nxz + g) : z4;
            if constexpr (SBP) {
                const float* C = a.curr + shot * a.shot_stride;
                qmx[m] = ok ? ld4(C + g) : z4;
                qmz[m] = ok ? ld4(C + nxz + g) : z4;
            } else {
                const long long fo = shot * a.fshot_stride + g;
                qmx[m] = ok && a.img_mode != 1 ? ld4(a.fm + fo) : z4;
                qmz[m] = ok && a.img_mode != 1 ? ld4(a.fm + fo + nxz) : z4;
                qpx[m] = ok && a.img_mode != 2 ? ld4(a.fp + fo) : z4;
                qpz[m] = ok && a.img_mode != 2 ? ld4(a.fp + fo + nxz) : z4;
            }
        }
        __builtin_amdgcn_s_waitcnt(0);  // the windows (and cell values) have landed
        __syncthreads();
        // the next shot's windows go into the other buffer while this one is imaged
        if (NBUF == 2 && shot + 1 < s1) issue_windows(shot + 1, b ^ 1);
        const float* sAx = sW + b * 4 * WS;
        const float* sAz = sAx + WS;
        const float* sFx = sAx + 2 * WS;
        const float* sFz = sAx + 3 * WS;
#pragma unroll
        for (int m = 0; m < NQ3; m++) {
            const int i = tid + m * kThreads;
            const int gx = sx + i / QC, z0 = sz + 4 * (i % QC);
            if (gx >= ex || z0 >= ez) continue;
            const int w = (gx - wsx) * WZ + (z0 - wsz);
            if constexpr (SBP) {
                // density term by parts: snapshot j against the weighted adjoint levels
                float f0x[4], f0z[4], px[4], pz[4], cx[4], cz[4], nx4[4], nz4[4];
                unpack(ld4(&sFx[w]), f0x);
                unpack(ld4(&sFz[w]), f0z);
                unpack(ld4(&sAx[w]), px);
                unpack(ld4(&sAz[w]), pz);
                unpack(qmx[m], cx);
                unpack(qmz[m], cz);
                unpack(qnx[m], nx4);
                unpack(qnz[m], nz4);
                const bool st = VTI && a.stiff;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (z0 + k >= nz) continue;
                    const float ax = a.wn * nx4[k] + a.wc * cx[k] + a.wp * px[k];
                    const float az = a.wn * nz4[k] + a.wc * cz[k] + a.wp * pz[k];
                    const float r = f0x[k] * ax + f0z[k] * az;
                    if (VTI) {
                        gr[m][k] += st ? 0.f : r;
                        g4[VTI ? m : 0][k] += st ? r : 0.f;
                    } else {
                        gr[m][k] += r;
                    }
                }
            } else {
                // density term first, so that the cell values are dead before the
                // derivative reads
                float f0x[4], f0z[4], fmx[4], fmz[4], fpx[4], fpz[4], ox[4], oz[4];
                unpack(ld4(&sFx[w]), f0x);
                unpack(ld4(&sFz[w]), f0z);
                unpack(qmx[m], fmx);
                unpack(qmz[m], fmz);
                unpack(qpx[m], fpx);
                unpack(qpz[m], fpz);
                unpack(qnx[m], ox);
                unpack(qnz[m], oz);
                image_rho<VTI>(a, z0, f0x, f0z, fmx, fmz, fpx, fpz, ox, oz, gr[m], g4[VTI ? m : 0]);
            }
            if (SBP && !a.img_deriv) continue;
            __builtin_amdgcn_sched_barrier(0);
            float axx[4], azx[4], axz[4], azz[4], t0x[4], t1z[4], t2x[4], t3z[4];
            if (inner) {
                quad_derivs<false, WZ>(a, sAx, sAz, w, gx, z0, axx, azx, axz, azz);
                quad_derivs<false, WZ>(a, sFx, sFz, w, gx, z0, t0x, t3z, t2x, t1z);
            } else {
                quad_derivs<true, WZ>(a, sAx, sAz, w, gx, z0, axx, azx, axz, azz);
                quad_derivs<true, WZ>(a, sFx, sFz, w, gx, z0, t0x, t3z, t2x, t1z);
            }
            const long long g = (long long)gx * ldz + z0;
            image_stiff<VTI>(a, g, z0, axx, azx, axz, azz, t0x, t1z, t2x, t3z, gl[m], gm[m], gr[m],
                             g3[VTI ? m : 0]);
        }
        if (NBUF == 2) __syncthreads();  // every wavefront is done with buffer b
    }
    if (s0 >= s1) return;
    float* G = a.grad + s0 * a.gslab_stride;
#pragma unroll
    for (int m = 0; m < NQ3; m++) {
        const int i = tid + m * kThreads;
        const int gx = sx + i / QC, z0 = sz + 4 * (i % QC);
        if (gx >= ex || z0 >= ez) continue;
        const long long g = (long long)gx * ldz + z0;
        float4 v = ld4(G + g);
        st4(G + g, make_float4(v.x + gl[m][0], v.y + gl[m][1], v.z + gl[m][2], v.w + gl[m][3]));
        v = ld4(G + nxz + g);
        st4(G + nxz + g, make_float4(v.x + gm[m][0], v.y + gm[m][1], v.z + gm[m][2], v.w + gm[m][3]));
        v = ld4(G + 2 * nxz + g);
        st4(G + 2 * nxz + g, make_float4(v.x + gr[m][0], v.y + gr[m][1], v.z + gr[m][2], v.w + gr[m][3]));
        if (VTI && a.stiff) {
            v = ld4(G + 3 * nxz + g);
            st4(G + 3 * nxz + g, make_float4(v.x + g3[m][0], v.y + g3[m][1], v.z + g3[m][2], v.w + g3[m][3]));
            v = ld4(G + 4 * nxz + g);
            st4(G + 4 * nxz + g, make_float4(v.x + g4[m][0], v.y + g4[m][1], v.z + g4[m][2], v.w + g4[m][3]));
        }
    }
}

// ---------------------------------------------------------------------------
// Initial condition: prev = 1/2 dt^2/rho * inject(trace[0])  (we_op.cpp:655-663)
// ---------------------------------------------------------------------------
__global__ void k_init_prev(float* prev, long long shot_stride, long long nxz, int ldz, const int* shot_ptr,
                            const PointEntry* ents, const float* trace)
{
    // one single-thread block per shot: deterministic accumulation order over its injectors
    const int s = blockIdx.x;
    const int b = shot_ptr[s], e = shot_ptr[s + 1];
    float* px = prev + s * shot_stride;
    float* pz = px + nxz;
    for (int k = b; k < e; k++) {
        const PointEntry en = ents[k];
        const float vx = trace[en.trace], vz = trace[en.tz];
        for (int ix = 0; ix < 3; ix++)
            for (int iz = 0; iz < 3; iz++) {
                const long long g = (long long)(en.xi0 + ix) * ldz + (en.zi0 + iz);
                const float w = en.wx[ix] * en.wz[iz];
                px[g] += w * vx;
                pz[g] += w * vz;
            }
    }
}

__global__ void k_scale_prev(float* prev, const float* rho, float half_dt2, long long nxz, int nshots)
{
    const long long n = 2 * nxz * nshots;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long g = i % nxz;
        prev[i] *= half_dt2 / rho[g];
    }
}

// extraction of one level from global memory (used for the last level)
__global__ void k_extract(const float* state, long long shot_stride, long long nxz, int ldz, const PointEntry* ents,
                          const int* rptr, int nrec, float* trace, int it)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * nrec) return;
    const int r = k >> 1, comp = k & 1;
    float v = 0.f;
    long long out = 0;
    for (int f = rptr[r]; f < rptr[r + 1]; f++) {
        const PointEntry en = ents[f];
        const float* u = state + en.shot * shot_stride + comp * nxz;
        for (int ix = 0; ix < 3; ix++)
            for (int iz = 0; iz < 3; iz++) v += en.wx[ix] * en.wz[iz] * u[(long long)(en.xi0 + ix) * ldz + en.zi0 + iz];
        out = comp ? en.tz : en.trace;
    }
    if (rptr[r + 1] > rptr[r]) trace[out + it] = v;
}

// ---------------------------------------------------------------------------
// Point sources of a step (we_op.cpp:722-724): added to the new level after the step,
// one thread per touched cell summing its contributions in a fixed order.  The step
// forms next = (dt^2/rho (r + s) + 2 curr - prev) / d * taper, which is linear in the
// source term s, so the sources add dt^2/rho * s / d * taper (d: the absorbing divisor on
// the perimeter, 1 elsewhere).
// ---------------------------------------------------------------------------
__global__ void k_inject_cells(const InjCell* cells, int ncell, const int* ptr, const long long* ctx,
                               const long long* ctz, const float* cw, const float* trace, int it, float* next,
                               long long shot_stride, long long nxz, int nx, int nz, int ldz, const float* irho,
                               const float* taper_x, const float* taper_z, const float* div_tb, const float* div_lr,
                               const InjImage im)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const InjCell cell = cells[c];
    float fx = 0.f, fz = 0.f;
    for (int k = ptr[c]; k < ptr[c + 1]; k++) {
        fx += cw[k] * trace[ctx[k] + it];
        fz += cw[k] * trace[ctz[k] + it];
    }
    const int gx = cell.gx, gz = cell.gz;
    const long long g = (long long)gx * ldz + gz;
    float dx = 1.f, dz = 1.f;
    if (gz == 0) { dx = div_tb[gx]; dz = div_tb[nx + gx]; }
    else if (gz == nz - 1) { dx = div_tb[2 * nx + gx]; dz = div_tb[3 * nx + gx]; }
    else if (gx == 0) { dx = div_lr[gz]; dz = div_lr[nz + gz]; }
    else if (gx == nx - 1) { dx = div_lr[2 * nz + gz]; dz = div_lr[3 * nz + gz]; }
    const float f = irho[g] * taper_z[gz] * taper_x[gx];
    float* u = next + cell.shot * shot_stride + g;
    const float ix = f * fx / dx, iz = f * fz / dz;
    u[0] += ix;
    u[nxz] += iz;
    if (im.on) {
        // the fused imaging step took the new level before this increment: add its share
        // of the density term by parts, wn f0 . (ix, iz), to the shot's own gradient slab
        const long long fo = cell.shot * im.fshot_stride + g;
        im.grad[cell.shot * im.gslab_stride + im.rho_off + g] += im.wn * (ix * im.f0[fo] + iz * im.f0[fo + nxz]);
    }
}

// ---------------------------------------------------------------------------
// Imaging condition (we_op.cpp:505-562 iso, 1478-1557 vti), one adjoint step
// ---------------------------------------------------------------------------
template <bool VTI, int TX>
__global__ void __launch_bounds__(kThreads) k_image(const ImageArgs a)
{
    constexpr int TZ = kTileZ, H = 2;
    constexpr int WZ = TZ + 2 * H, WX = TX + 2 * H;
    __shared__ float sAx[WX * WZ], sAz[WX * WZ], sFx[WX * WZ], sFz[WX * WZ];
    const int ntiles = a.ntx * a.ntz;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int shot = lin / ntiles;
    const int tile = lin - shot * ntiles;
    const int tix = tile / a.ntz, tiz = tile - (tile / a.ntz) * a.ntz;
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    int sx, ex, sz, ez;
    tile_range_x_hd(tix, a.ntx, TX, nx, sx, ex);
    tile_range_z_hd(tiz, a.ntz, nz, ldz, sz, ez);
    const int wsx = window_start(sx, H), wsz = window_start(sz, H);
    const long long nxz = (long long)nx * ldz;
    const float* dA = (a.lag ? a.adj_prev : a.adj) + shot * a.shot_stride;
    const float* F0 = a.f0 + shot * a.fshot_stride;
    for (int i = threadIdx.x; i < WX * WZ; i += kThreads) {
        const int lx = i / WZ, lz = i - lx * WZ;
        const int gx = wsx + lx, gz = wsz + lz;
        float ax = 0.f, az = 0.f, fx = 0.f, fz = 0.f;
        if (gx >= 0 && gx < nx && gz >= 0 && gz < nz) {
            const long long g = (long long)gx * ldz + gz;
            ax = dA[g];
            az = dA[nxz + g];
            fx = F0[g];
            fz = F0[nxz + g];
        }
        sAx[i] = ax;
        sAz[i] = az;
        sFx[i] = fx;
        sFz[i] = fz;
    }
    __syncthreads();
    const float* A = a.adj + shot * a.shot_stride;
    const float* Fm = a.fm ? a.fm + shot * a.fshot_stride : nullptr;
    const float* Fp = a.fp ? a.fp + shot * a.fshot_stride : nullptr;
    float* G = a.grad + shot * a.gshot_stride;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int gz = sz + lane; gz < ez; gz += 64) {
        for (int j = 0; j < TX / 4; j++) {
            const int gx = sx + wave * (TX / 4) + j;
            if (gx >= ex) break;
            const int w = (gx - wsx) * WZ + (gz - wsz);
            const long long g = (long long)gx * ldz + gz;
            const float axx = d1<false>(&sAx[w], WZ, gx, nx) * a.rdx;  // adjoint x_x
            const float azx = d1<false>(&sAx[w], 1, gz, nz) * a.rdz;   // adjoint x_z
            const float axz = d1<false>(&sAz[w], WZ, gx, nx) * a.rdx;  // adjoint z_x
            const float azz = d1<false>(&sAz[w], 1, gz, nz) * a.rdz;   // adjoint z_z
            const float t0 = d1<false>(&sFx[w], WZ, gx, nx) * a.rdx;   // forward x_x
            const float t1 = d1<false>(&sFz[w], 1, gz, nz) * a.rdz;    // forward z_z
            const float t2 = d1<false>(&sFz[w], WZ, gx, nx) * a.rdx;   // forward z_x
            const float t3 = d1<false>(&sFx[w], 1, gz, nz) * a.rdz;    // forward x_z
            const float f0x = F0[g], f0z = F0[nxz + g];
            float d2x, d2z;
            if (a.mode == 0) {
                d2x = Fp[g] - 2.f * f0x + Fm[g];
                d2z = Fp[nxz + g] - 2.f * f0z + Fm[nxz + g];
            } else if (a.mode == 1) {
                d2x = Fp[g] - f0x;
                d2z = Fp[nxz + g] - f0z;
            } else {
                d2x = -f0x + Fm[g];
                d2z = -f0z + Fm[nxz + g];
            }
            if (!VTI) {
                G[g] += a.wlm * (axx + azz) * (t0 + t1);
                G[nxz + g] += a.wlm * ((azx + axz) * (t2 + t3) + 2.f * axx * t0 + 2.f * azz * t1);
            } else {
                const float l = a.lam[g], m = a.mu[g], c = a.c13[g];
                const float del = ((c + m) * (c + m) - (l + m) * (l + m)) / (2.f * (l + m) * (l + 2.f * m));
                const float v1 = sqrtf(2.f * (l + 2.f * m) * (l + m) * del + (l + m) * (l + m));
                const float v2 = ((1.f + 2.f * del) * l + (1.f + 3.f * del) * m) / v1;
                const float v3 = ((1.f + 3.f * del) * l + (1.f + 4.f * del) * m) / v1 - 1.f;
                const float e = 1.f + 2.f * a.eps[g];
                G[g] += a.wlm * (e * axx * t0 + azz * t1 + v2 * (axx * t1 + azz * t0));
                G[nxz + g] += a.wlm * ((azx + axz) * (t2 + t3) + 2.f * e * axx * t0 + 2.f * azz * t1 +
                                       v3 * (axx * t1 + azz * t0));
            }
            G[2 * nxz + g] += a.wrho * (A[g] * d2x + A[nxz + g] * d2z);
        }
    }
}

// ---------------------------------------------------------------------------
// Utilities
// ---------------------------------------------------------------------------
__global__ void k_sum_shots(const float* per_shot, long long shot_stride, int nshots, long long n, float* out, int acc)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float v = acc ? out[i] : 0.f;
        for (int s = 0; s < nshots; s++) v += per_shot[s * shot_stride + i];
        out[i] = v;
    }
}

__global__ void k_fill(float* p, long long n, float v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// dipole_to_strain (misc.cpp:18-43) of DAS channels, in place on traces [2][nr][nt]:
// forward x = (c x + s z) / gl, z = 0; adjoint z = s x / gl, x = c x / gl, with cs = (cos,
// sin) of each channel's dip; in double precision, one sample per thread.
__global__ void k_das_strain(float* tr, const double* cs, int nr, int nt, double gl, int adjoint)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nr * nt) return;
    const int r = (int)(i / nt);
    const double c = cs[2 * r], sn = cs[2 * r + 1];
    float* px = tr + i;
    float* pz = tr + (long long)nr * nt + i;
    const double x = *px;
    if (!adjoint) {
        *px = (float)((c * x + sn * (double)*pz) / gl);
        *pz = 0.f;
    } else {
        *pz = (float)(sn * x / gl);
        *px = (float)(c * x / gl);
    }
}

// The per-trace time map of the operator level (resampling, time derivative, quadrature
// weights, reversal, composed on the host): out[x][i] = sum_k w[k] in[x][col[k]] over the
// row-compressed entries k of row i, accumulated in double precision; one output sample
// per thread, consecutive threads on consecutive samples of one trace.
__global__ void k_trace_matvec(const int* __restrict__ ptr, const int* __restrict__ col, const double* __restrict__ w,
                               int nrow, int ncol, const float* __restrict__ in, float* __restrict__ out, long long n)
{
    for (long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x; gid < n;
         gid += (long long)gridDim.x * blockDim.x) {
        const long long x = gid / nrow;
        const int i = (int)(gid - x * nrow);
        const float* a = in + x * ncol;
        double acc = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) acc += w[k] * (double)a[col[k]];
        out[gid] = (float)acc;
    }
}

// 16-bit checkpoint levels (BF16 or FP16) with a max-abs scale per (level, shot):
// x = round16(u / s), u = s x.
// Scales are kept as the bit patterns of non-negative floats, whose integer order is the
// float order, so the per-shot maximum is an integer atomic max.
__global__ void k_absmax_shots(const float* u, long long per_shot, int nshots, unsigned* amax)
{
    const int shot = blockIdx.y;
    const float* p = u + shot * per_shot;
    float m = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per_shot;
         i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(p[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(&amax[shot], __float_as_uint(m));
}

__device__ __forceinline__ unsigned short to_bf16(float x)
{
    const unsigned b = __float_as_uint(x);
    return (unsigned short)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);  // round to nearest even
}

template <bool F16>
__global__ void k_pack16(const float* u, unsigned short* q, const unsigned* amax, float* scale, long long per_shot)
{
    const int shot = blockIdx.y;
    const float s = __uint_as_float(amax[shot]);
    const float inv = s > 0.f ? 1.f / s : 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) scale[shot] = s;
    const float* p = u + shot * per_shot;
    unsigned short* d = q + shot * per_shot;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per_shot;
         i += (long long)gridDim.x * blockDim.x)
        d[i] = F16 ? __half_as_ushort(__float2half_rn(p[i] * inv)) : to_bf16(p[i] * inv);
}

template <bool F16>
__global__ void k_unpack16(const unsigned short* q, const float* scale, float* u, long long per_shot)
{
    const int shot = blockIdx.y;
    const float s = scale[shot];
    const unsigned short* p = q + shot * per_shot;
    float* d = u + shot * per_shot;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < per_shot;
         i += (long long)gridDim.x * blockDim.x)
        d[i] = (F16 ? __half2float(__ushort_as_half(p[i])) : __uint_as_float((unsigned)p[i] << 16)) * s;
}

__global__ void k_copy(float* dst, const float* src, long long n)
{
    const long long n4 = n >> 2;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) d4[i] = s4[i];
    for (long long i = (n4 << 2) + blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) dst[i] = src[i];
}

// partial[b] = sum over this block's elements of a*b*w (b, w optional) in FP64
__global__ void k_dot_f64(const float* x, const float* y, const float* w, long long n, double* partial)
{
    __shared__ double red[kThreads];
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double v = (double)x[i] * (y ? (double)y[i] : (double)x[i]);
        if (w) v *= (double)w[i];
        acc += v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_reduce_f64(const double* partial, int n, double* out)
{
    __shared__ double red[kThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kThreads) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int grid_for(long long n, int cap = 4096)
{
    long long b = (n + kThreads - 1) / kThreads;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

template <bool VTI, int TX, int TZ, bool EDGE>
static void step_launch(const StepArgs& a, const int4* boxes, const int* list, int n, bool image, hipStream_t s)
{
    const int nb = n * a.ngroups;
    if (nb <= 0) return;
    if (image) hipLaunchKernelGGL((k_step<VTI, TX, TZ, EDGE, true>), dim3(nb), dim3(kThreads), 0, s, a, boxes, list);
    else hipLaunchKernelGGL((k_step<VTI, TX, TZ, EDGE, false>), dim3(nb), dim3(kThreads), 0, s, a, boxes, list);
}

template <bool VTI>
static void step_class(const StepArgs& a, BoxClass c, const int4* boxes, const int* list, int n, bool image,
                       hipStream_t s)
{
    switch (c) {
    case BoxClass::core16: step_launch<VTI, 16, 64, false>(a, boxes, list, n, image, s); break;
    case BoxClass::core32: step_launch<VTI, 32, 64, false>(a, boxes, list, n, image, s); break;
    case BoxClass::band_z: step_launch<VTI, 64, 16, true>(a, boxes, list, n, image, s); break;
    case BoxClass::band_z8: step_launch<VTI, kBandZ8Width, 8, true>(a, boxes, list, n, image, s); break;
    case BoxClass::band_x: step_launch<VTI, 8, 128, true>(a, boxes, list, n, image, s); break;
    }
}

void launch_step(const StepArgs& a, bool vti, BoxClass c, const int4* boxes, const int* list, int n, bool image,
                 hipStream_t s)
{
    if (vti) step_class<true>(a, c, boxes, list, n, image, s);
    else step_class<false>(a, c, boxes, list, n, image, s);
}

template <bool VTI, int NBUF, bool SBP>
static void img_dispatch(const StepArgs& a, int tile_x, int n, hipStream_t s)
{
    if (tile_x == 32) k_img<VTI, 32, NBUF, SBP><<<n, kThreads, 0, s>>>(a);
    else k_img<VTI, 16, NBUF, SBP><<<n, kThreads, 0, s>>>(a);
}

template <bool VTI, bool SBP>
static void img_bufs(const StepArgs& a, int tile_x, int nbuf, int n, hipStream_t s)
{
    if (nbuf == 2) img_dispatch<VTI, 2, SBP>(a, tile_x, n, s);
    else img_dispatch<VTI, 1, SBP>(a, tile_x, n, s);
}

void launch_img(const StepArgs& a, bool vti, int tile_x, int nbuf, bool sbp, hipStream_t s)
{
    const int n = a.ntx * a.ntz * a.ngroups;
    if (n <= 0) return;
    if (vti) {
        if (sbp) img_bufs<true, true>(a, tile_x, nbuf, n, s);
        else img_bufs<true, false>(a, tile_x, nbuf, n, s);
    } else {
        if (sbp) img_bufs<false, true>(a, tile_x, nbuf, n, s);
        else img_bufs<false, false>(a, tile_x, nbuf, n, s);
    }
}

template <bool VTI>
static void image_dispatch(const ImageArgs& a, int tile_x, hipStream_t s)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    if (tile_x == 32) hipLaunchKernelGGL((k_image<VTI, 32>), dim3(nb), dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((k_image<VTI, 16>), dim3(nb), dim3(kThreads), 0, s, a);
}

void launch_image(const ImageArgs& a, bool vti, int tile_x, hipStream_t s)
{
    if (vti) image_dispatch<true>(a, tile_x, s);
    else image_dispatch<false>(a, tile_x, s);
}

void launch_init_prev(float* prev, long long shot_stride, long long nxz, int ldz, int nshots, const int* shot_ptr,
                      const PointEntry* ents, const float* trace, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_prev, dim3(nshots), dim3(1), 0, s, prev, shot_stride, nxz, ldz, shot_ptr, ents, trace);
}

void launch_scale_prev(float* prev, const float* rho, float half_dt2, long long nxz, int nshots, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_prev, dim3(grid_for(2 * nxz * nshots)), dim3(kThreads), 0, s, prev, rho, half_dt2, nxz,
                       nshots);
}

void launch_inject_cells(const InjCell* cells, int ncell, const int* ptr, const long long* ctx, const long long* ctz,
                         const float* cw, const float* trace, int it, float* next, long long shot_stride,
                         long long nxz, int nx, int nz, int ldz, const float* irho, const float* taper_x,
                         const float* taper_z, const float* div_tb, const float* div_lr, const InjImage& im,
                         hipStream_t s)
{
    if (ncell <= 0) return;
    hipLaunchKernelGGL(k_inject_cells, dim3((ncell + kThreads - 1) / kThreads), dim3(kThreads), 0, s, cells, ncell,
                       ptr, ctx, ctz, cw, trace, it, next, shot_stride, nxz, nx, nz, ldz, irho, taper_x,
                       taper_z, div_tb, div_lr, im);
}

void launch_extract(const float* state, long long shot_stride, long long nxz, int ldz, const PointEntry* ents,
                    const int* rptr, int nrec, float* trace, int it, hipStream_t s)
{
    if (nrec <= 0) return;
    hipLaunchKernelGGL(k_extract, dim3((2 * nrec + kThreads - 1) / kThreads), dim3(kThreads), 0, s, state,
                       shot_stride, nxz, ldz, ents, rptr, nrec, trace, it);
}

void launch_sum_shots(const float* per_shot, long long shot_stride, int nshots, long long n, float* out, int acc,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_shots, dim3(grid_for(n)), dim3(kThreads), 0, s, per_shot, shot_stride, nshots, n, out, acc);
}

void launch_fill(float* p, long long n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
}

void launch_trace_matvec(const int* ptr, const int* col, const double* w, int nrow, int ncol, const float* in,
                         float* out, long long ntr, hipStream_t s)
{
    const long long n = ntr * nrow;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_trace_matvec, dim3(grid_for(n)), dim3(kThreads), 0, s, ptr, col, w, nrow, ncol, in, out, n);
}

void launch_das_strain(float* tr, const double* cs, int nr, int nt, double gl, bool adjoint, hipStream_t s)
{
    const long long n = (long long)nr * nt;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_das_strain, dim3(grid_for(n)), dim3(kThreads), 0, s, tr, cs, nr, nt, gl, adjoint ? 1 : 0);
}

void launch_pack16(bool f16, const float* u, unsigned short* q, float* scale, unsigned* amax, long long per_shot,
                   int nshots, hipStream_t s)
{
    const int gx = (int)std::min<long long>((per_shot + kThreads - 1) / kThreads, 512);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kThreads), 0, s, (float*)amax, (long long)nshots, 0.f);
    hipLaunchKernelGGL(k_absmax_shots, dim3(gx, nshots), dim3(kThreads), 0, s, u, per_shot, nshots, amax);
    if (f16) hipLaunchKernelGGL(k_pack16<true>, dim3(gx, nshots), dim3(kThreads), 0, s, u, q, amax, scale, per_shot);
    else hipLaunchKernelGGL(k_pack16<false>, dim3(gx, nshots), dim3(kThreads), 0, s, u, q, amax, scale, per_shot);
}

void launch_unpack16(bool f16, const unsigned short* q, const float* scale, float* u, long long per_shot, int nshots,
                     hipStream_t s)
{
    const int gx = (int)std::min<long long>((per_shot + kThreads - 1) / kThreads, 512);
    if (f16) hipLaunchKernelGGL(k_unpack16<true>, dim3(gx, nshots), dim3(kThreads), 0, s, q, scale, u, per_shot);
    else hipLaunchKernelGGL(k_unpack16<false>, dim3(gx, nshots), dim3(kThreads), 0, s, q, scale, u, per_shot);
}

void launch_copy(float* dst, const float* src, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3(grid_for((n + 3) / 4)), dim3(kThreads), 0, s, dst, src, n);
}

void launch_dot_f64(const float* x, const float* y, const float* w, long long n, double* partial, int nblocks,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_f64, dim3(nblocks), dim3(kThreads), 0, s, x, y, w, n, partial);
}

void launch_reduce_f64(const double* partial, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_f64, dim3(1), dim3(kThreads), 0, s, partial, n, out);
}

// ---------------------------------------------------------------------------
// Acoustic operator (nl_we_op_a::propagate, we_op.cpp:1937-2135; the SAT terms and their
// scaling, spatial_operators.cpp:551-860).  A workgroup owns a 32 (x) x 32 (z) tile of
// a group of shots; the level and 1/rho windows go through LDS (8 rows above and below
// the tile, which holds every closure row a tile can need; the columns as below), and
// each thread forms the new level of one quad (column-major, so a wavefront covers 8
// whole columns): D2x + D2z of the 1/rho-weighted Laplacian (interior weights or closure
// rows per axis), the SAT terms of the boundary cells, the leapfrog with dt^2 K, the
// absorbing boundary divisor and the cos^2 taper.  The reference tapers the current
// level in place after each step; here the previous level is tapered on load instead
// (taper_prev), which gives the same values.
// ---------------------------------------------------------------------------
constexpr int kAcH = 8;
constexpr int kAcWZ = kAcTZ + 2 * kAcH, kAcWQ = kAcWZ / 4;
static_assert(kAcTX * (kAcTZ / 4) == kThreads, "one quad per thread");

// SAT terms of one side at a cell k cells in from it (0 <= k < 4), added to rhs: pu / pc
// point at the side's boundary cell of the cell's line in the level / 1/rho windows, st
// the window stride pointing inwards, p the cell's previous level (absorbing), K the
// boundary cell's bulk modulus (absorbing).
__device__ __forceinline__ float ac_sat(int kind, int k, const float* pu, const float* pc, int st, float p, float K,
                                        const float* dir, float pen, float neu, float absc)
{
    if (kind == 1) {
        // Dirichlet (free surface): + H^-1 S^T (ri u)_0 on the first 4 cells, - penalty at 0
        float v = dir[k] * pc[0] * pu[0];
        if (k == 0) v -= pc[0] * pu[0] * pen;
        return v;
    }
    if (k != 0) return 0.f;
    // Neumann: - H^-1 (ri S u)_0; absorbing: + the impedance term of the previous level
    float v = -neu * pc[0] * bnd_deriv(pu, st);
    if (kind == 2) v += absc * p / sqrtf(K / pc[0]);
    return v;
}

// The step's level window is narrower in x than the imaging windows: 2 columns before the
// tile and 6 after it, the window moved left to start at column nx - 8 at the far side.
// That holds every point a tile's cells read: interior D2 rows +-2, the closure rows of
// columns 0-5 read columns 0-7 (first tile) and those of columns nx-6..nx-1 columns
// nx-8..nx-1 (a tile holding such a column ends at most 5 columns before its window does),
// and the SAT terms the boundary column and its 3 neighbours.
constexpr int kAcSHL = 2, kAcSWX = kAcTX + kAcSHL + 6;

__device__ __forceinline__ float sel4(const float (&v)[4], int k)
{
    return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}

__device__ __forceinline__ void ac_window_x(const float* u, float* sU, int wsx, int wsz, int nx, int ldz)
{
    for (int i = threadIdx.x; i < kAcSWX * kAcWQ; i += kThreads) {
        const int lx = i / kAcWQ, lq = i - lx * kAcWQ;
        const int gx = wsx + lx, gz = wsz + 4 * lq;
        const bool ok = gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
        st4(&sU[lx * kAcWZ + 4 * lq], ok ? ld4(u + (long long)gx * ldz + gz) : make_float4(0.f, 0.f, 0.f, 0.f));
    }
}

// The new level of one frame cell with closure rows or SAT terms (the reference's
// Dxx_var/Dzz_var closures and asat_* terms, then the leapfrog, the boundary divisor and
// the taper), from the level window U and the 1/rho window R (w: the cell's window index)
__device__ __forceinline__ float ac_frame_cell(const AcousticArgs& a, const float* U, const float* R, int w, int wsx,
                                               int wsz, int gx, int gz, long long so)
{
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    const long long gi = (long long)gx * ldz + gz;
    const float* pu = &U[w];
    const float* pc = &R[w];
    const float tzk = a.taper_z[gz], tx = a.taper_x[gx];
    float p = a.prev[so + gi];
    if (a.taper_prev) p = (p * tzk) * tx;
    const float lx = (gx >= 6 && gx < nx - 6) ? d2_apply(d2_weights(pc, kAcWZ), pu, kAcWZ)
                                              : d2_closure(pu, pc, kAcWZ, gx, nx);
    const float lz = (gz >= 6 && gz < nz - 6) ? d2_apply(d2_weights(pc, 1), pu, 1) : d2_closure(pu, pc, 1, gz, nz);
    float rhs = lx * a.rdx2 + lz * a.rdz2;
    const int col = (gx - wsx) * kAcWZ;
    if (gz < 4) {  // top side, boundary cell (gx, 0)
        const int bb = col - wsz;
        rhs += ac_sat(a.bc[0], gz, &U[bb], &R[bb], 1, p, gz == 0 ? a.K[gi - gz] : 0.f, a.dir_z, a.pen_z, a.neu_z,
                      a.abs_z);
    }
    if (gz >= nz - 4) {  // bottom side, boundary cell (gx, nz - 1)
        const int bb = col + (nz - 1 - wsz);
        rhs += ac_sat(a.bc[1], nz - 1 - gz, &U[bb], &R[bb], -1, p, gz == nz - 1 ? a.K[gi] : 0.f, a.dir_z, a.pen_z,
                      a.neu_z, a.abs_z);
    }
    if (gx < 4) {  // left side, boundary cell (0, gz)
        const int bb = (0 - wsx) * kAcWZ + (gz - wsz);
        rhs += ac_sat(a.bc[2], gx, &U[bb], &R[bb], kAcWZ, p, gx == 0 ? a.K[gi] : 0.f, a.dir_x, a.pen_x, a.neu_x,
                      a.abs_x);
    }
    if (gx >= nx - 4) {  // right side, boundary cell (nx - 1, gz)
        const int bb = (nx - 1 - wsx) * kAcWZ + (gz - wsz);
        rhs += ac_sat(a.bc[3], nx - 1 - gx, &U[bb], &R[bb], -kAcWZ, p, gx == nx - 1 ? a.K[gi] : 0.f, a.dir_x,
                      a.pen_x, a.neu_x, a.abs_x);
    }
    // leapfrog (we_op.cpp:2093-2098), boundary divisor (asat_scale_boundaries), taper
    const float v = (a.kdt2[gi] * rhs + (2.f * pu[0] - p)) / a.div[gi];
    return (v * tzk) * tx;
}

// A workgroup steps one 32 x 32 tile of a.spb consecutive shots: the 1/rho window, dt^2 K,
// the boundary divisor and the tapers are loaded once (and, for tiles of interior cells
// only, the D2 weights of the cells formed once); the shots' level windows alternate
// between two LDS buffers, the next shot's loaded while the current one is computed (one
// barrier per shot).  EDGE = false: a tile of the rectangle of interior tiles; EDGE = true:
// a tile of the frame around it, whose interior cells take the interior rows (weights
// formed per shot) and whose other cells the closure rows and SAT terms, dealt one cell
// per thread.
// IMG: the adjoint step also images snapshot a.f0 of every shot against the level it
// reads (the curr window is the adjoint level compute_gradients takes), as k_agrad does,
// with the snapshot's window double buffered beside the level's.
template <bool EDGE, bool IMG>
__device__ __forceinline__ void astep_tile(const AcousticArgs& a, int tix, int tiz, int grp, float* sR,
                                           float (*sU)[kAcSWX * kAcWZ], float* sO, float (*sF)[kAcSWX * kAcWZ])
{
    const int sx = tix * kAcTX, sz = tiz * kAcTZ;
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    const int ex = min(sx + kAcTX, nx), ez = min(sz + kAcTZ, ldz);
    const int wsx = min(sx - kAcSHL, nx - 8), wsz = sz - kAcH;
    const int s0 = grp * a.spb, s1 = min(s0 + a.spb, a.nshots);
    for (int i = threadIdx.x; i < kAcSWX * kAcWQ; i += kThreads) {
        const int lx = i / kAcWQ, lq = i - lx * kAcWQ;
        const int gx = wsx + lx, gz = wsz + 4 * lq;
        const bool ok = gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
        st4(&sR[lx * kAcWZ + 4 * lq], ok ? ld4(a.ri + (long long)gx * ldz + gz) : make_float4(0.f, 0.f, 0.f, 0.f));
    }
    ac_window_x(a.curr + (long long)s0 * a.shot_stride, sU[0], wsx, wsz, nx, ldz);
    if (IMG) ac_window_x(a.f0 + (long long)s0 * a.fshot_stride, sF[0], wsx, wsz, nx, ldz);
    const int q = threadIdx.x;  // quads column-major: a wavefront covers 8 whole columns
    const int col_q = q / (kAcTZ / 4), row_q = q % (kAcTZ / 4);
    const int gx = sx + col_q, z0 = sz + 4 * row_q;
    const bool active = gx < ex && z0 < ez;
    const long long g = active ? (long long)gx * ldz + z0 : 0;
    float kd[4], dv[4], tz[4];
    unpack(ld4(a.kdt2 + g), kd);
    unpack(ld4(a.div + g), dv);
    unpack(ld4(a.taper_z + (active ? z0 : 0)), tz);
    const float tx = a.taper_x[active ? gx : 0];
    const int col = (active ? gx - wsx : 2) * kAcWZ;  // an in-window column either way
    const int zw = (active ? z0 : sz) - wsz;
    __syncthreads();
    // interior D2 weights of the quad's cells, once for all shots (interior tiles; a frame
    // tile forms them per shot: held across the shot loop they cost a wavefront per SIMD)
    W5 wx[4], wz[4];
    if (!EDGE) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            wx[k] = d2_weights(&sR[col + zw + k], kAcWZ);
            wz[k] = d2_weights(&sR[col + zw + k], 1);
        }
    }
    const bool inx = gx >= 6 && gx < nx - 6;
    // frame tiles: the cells with closure rows or SAT terms (the tile's columns within 6 of
    // a side, then its rows within 6 of a side in the other columns), dealt to all the
    // workgroup's threads one cell each, consecutive threads along a row or down a column
    const int zr = min(ez, nz), nrr = zr - sz, ncx = ex - sx;
    const int cl = sx < 6 ? min(6, ex) - sx : 0, crs = max(nx - 6, sx + cl), cr = max(0, ex - crs);
    const int rt = sz < 6 ? max(0, min(6, zr) - sz) : 0, rbs = max(nz - 6, sz + rt), rb = max(0, zr - rbs);
    const int nsc = cl + cr, nsr = rt + rb, nnc = ncx - nsc;
    const int nslow = EDGE ? nsc * nrr + nnc * nsr : 0;
    // imaging: the group's gradient slab and the K term's weights w / (dtg K^2)
    float gk[4] = {0.f, 0.f, 0.f, 0.f}, gr[4] = {0.f, 0.f, 0.f, 0.f}, kw[4] = {0.f, 0.f, 0.f, 0.f};
    float* G = a.grad + (long long)grp * a.gslab_stride + g;
    if (IMG && active) {
        float K[4];
        unpack(ld4(a.K + g), K);
#pragma unroll
        for (int k = 0; k < 4; k++) kw[k] = a.wk / (K[k] * K[k]);
        unpack(ld4(G), gk);
        unpack(ld4(G + a.shot_stride), gr);
    }
    for (int sh = s0; sh < s1; sh++) {
        const int b = (sh - s0) & 1;
        const long long so = (long long)sh * a.shot_stride;
        if (sh + 1 < s1) {
            ac_window_x(a.curr + so + a.shot_stride, sU[b ^ 1], wsx, wsz, nx, ldz);
            if (IMG) ac_window_x(a.f0 + (long long)(sh + 1) * a.fshot_stride, sF[b ^ 1], wsx, wsz, nx, ldz);
        }
        if (active) {
            float pv[4], out[4];
            unpack(ld4(a.prev + so + g), pv);
            const float* U = sU[b];
            if (!EDGE) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float* pu = &U[col + zw + k];
                    float p = pv[k];
                    if (a.taper_prev) p = (p * tz[k]) * tx;
                    const float rhs = d2_apply(wx[k], pu, kAcWZ) * a.rdx2 + d2_apply(wz[k], pu, 1) * a.rdz2;
                    // leapfrog (we_op.cpp:2093-2098), boundary divisor, taper
                    const float v = (kd[k] * rhs + (2.f * pu[0] - p)) / dv[k];
                    out[k] = (v * tz[k]) * tx;
                }
            } else {
                // the interior cells of the quad (weights formed per shot) and the rows
                // past nz (zero) into the output tile; the other cells follow below
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int gz = z0 + k;
                    const float* pu = &U[col + zw + k];
                    const float* pc = &sR[col + zw + k];
                    if (gz >= nz) {
                        sO[(gx - sx) * kAcTZ + (gz - sz)] = 0.f;  // rows past nz stay zero
                    } else if (inx && gz >= 6 && gz < nz - 6) {
                        float p = pv[k];
                        if (a.taper_prev) p = (p * tz[k]) * tx;
                        const float rhs = d2_apply(d2_weights(pc, kAcWZ), pu, kAcWZ) * a.rdx2 +
                                          d2_apply(d2_weights(pc, 1), pu, 1) * a.rdz2;
                        const float v = (kd[k] * rhs + (2.f * pu[0] - p)) / dv[k];
                        sO[(gx - sx) * kAcTZ + (gz - sz)] = (v * tz[k]) * tx;
                    }
                }
            }
            if (!EDGE) st4(a.next + so + g, make_float4(out[0], out[1], out[2], out[3]));
            if (IMG) {
                // compute_gradients of this shot (we_op.cpp:1883-1935), as k_agrad
                const long long fo = (long long)sh * a.fshot_stride;
                float fm[4] = {0.f, 0.f, 0.f, 0.f}, fp[4] = {0.f, 0.f, 0.f, 0.f};
                if (a.mode != 1) unpack(ld4(a.fm + fo + g), fm);
                if (a.mode != 2) unpack(ld4(a.fp + fo + g), fp);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int gz = z0 + k;
                    if (gz >= nz) continue;
                    const float* pa = &U[col + zw + k];
                    const float* pf = &sF[b][col + zw + k];
                    const float ax = d1<false>(pa, kAcWZ, gx, nx) * a.rdx, az = d1<false>(pa, 1, gz, nz) * a.rdz;
                    const float fx = d1<false>(pf, kAcWZ, gx, nx) * a.rdx, fz = d1<false>(pf, 1, gz, nz) * a.rdz;
                    const float d2 =
                        a.mode == 0 ? fp[k] - 2.f * pf[0] + fm[k] : (a.mode == 1 ? fp[k] - pf[0] : -pf[0] + fm[k]);
                    gk[k] -= kw[k] * pa[0] * d2;
                    gr[k] += a.wr * (ax * fx + az * fz);
                }
            }
        }
        if (EDGE) {
#pragma unroll 1
            for (int c = threadIdx.x; c < nslow; c += kThreads) {
                int cx, cz;
                if (c < nsc * nrr) {
                    const int j = c / nrr;
                    cz = sz + (c - j * nrr);
                    cx = j < cl ? sx + j : crs + (j - cl);
                } else {
                    const int c2 = c - nsc * nrr, r = c2 / nnc;
                    cx = sx + cl + (c2 - r * nnc);
                    cz = r < rt ? sz + r : rbs + (r - rt);
                }
                sO[(cx - sx) * kAcTZ + (cz - sz)] =
                    ac_frame_cell(a, sU[b], sR, (cx - wsx) * kAcWZ + (cz - wsz), wsx, wsz, cx, cz, so);
            }
            __syncthreads();  // the output tile is complete
            if (active) st4(a.next + so + g, ld4(&sO[(gx - sx) * kAcTZ + (z0 - sz)]));
        }
        __syncthreads();  // the next shot's window is in; this shot's window is read
    }
    if (IMG && active) {
        st4(G, make_float4(gk[0], gk[1], gk[2], gk[3]));
        st4(G + a.shot_stride, make_float4(gr[0], gr[1], gr[2], gr[3]));
    }
}

// One launch for all tiles: the frame tiles' workgroups first (their closure rows make
// them the slowest), then the interior rectangle's, each class spread over all XCDs (the
// frame's block count rounded up to a multiple of 8, the extra blocks idle) and remapped
// within its class so that an XCD runs neighbouring tiles.
template <bool IMG>
__global__ void __launch_bounds__(kThreads) k_astep(const AcousticArgs a)
{
    __shared__ __attribute__((aligned(16))) float sR[kAcSWX * kAcWZ];
    __shared__ __attribute__((aligned(16))) float sU[2][kAcSWX * kAcWZ];
    __shared__ __attribute__((aligned(16))) float sO[kAcTX * kAcTZ];  // a frame tile's new level
    __shared__ __attribute__((aligned(16))) float sF[IMG ? 2 : 1][IMG ? kAcSWX * kAcWZ : 4];  // snapshot windows
    const int nin_x = a.itx1 - a.itx0, nin_z = a.itz1 - a.itz0;
    const int nin = nin_x * nin_z, nedge = a.ntx * a.ntz - nin;
    const int groups = (a.nshots + a.spb - 1) / a.spb;
    const int ne = nedge * groups, ne8 = (ne + 7) & ~7;
    const int b = blockIdx.x;
    if (b < ne8) {
        if (b >= ne) return;  // padding: the whole workgroup leaves before any barrier
        // frame: the rows of tiles above and below the rectangle, then its side columns
        const int lin = xcd_remap(b, ne);
        const int grp = lin / nedge, tile = lin - grp * nedge;
        const int top = a.itz0 * a.ntx, bot = (a.ntz - a.itz1) * a.ntx;
        int tix, tiz;
        if (tile < top + bot) {
            const int r = tile < top ? tile : tile - top;
            tix = r % a.ntx;
            tiz = (tile < top ? 0 : a.itz1) + r / a.ntx;
        } else {
            const int r = tile - top - bot, w = a.ntx - nin_x;
            const int c = r % w;
            tiz = a.itz0 + r / w;
            tix = c < a.itx0 ? c : a.itx1 + (c - a.itx0);
        }
        astep_tile<true, IMG>(a, tix, tiz, grp, sR, sU, sO, reinterpret_cast<float (*)[kAcSWX * kAcWZ]>(&sF[0][0]));
    } else {
        const int lin = xcd_remap(b - ne8, gridDim.x - ne8);
        const int grp = lin / nin, tile = lin - grp * nin;
        astep_tile<false, IMG>(a, a.itx0 + tile / nin_z, a.itz0 + tile % nin_z, grp, sR, sU, sO,
                               reinterpret_cast<float (*)[kAcSWX * kAcWZ]>(&sF[0][0]));
    }
}

// nl_we_op_a::compute_gradients (we_op.cpp:1883-1935) for every shot: the K term from the
// second time difference of the snapshots against the adjoint level, the 1/rho term
// from first derivatives of both.  A workgroup images one tile of a.spb consecutive
// shots (the step kernel's groups and windows, adjoint level and snapshot windows double
// buffered), sums their terms in shot order in registers and adds the sum to the group's
// gradient slab once.
__global__ void __launch_bounds__(kThreads) k_agrad(const AcousticArgs a)
{
    __shared__ __attribute__((aligned(16))) float sA[2][kAcSWX * kAcWZ];
    __shared__ __attribute__((aligned(16))) float sF[2][kAcSWX * kAcWZ];
    const int ntile = a.ntx * a.ntz;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int grp = lin / ntile, tile = lin - grp * ntile;
    const int tix = tile / a.ntz, tiz = tile - tix * a.ntz;
    const int sx = tix * kAcTX, sz = tiz * kAcTZ;
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    const int ex = min(sx + kAcTX, nx), ez = min(sz + kAcTZ, ldz);
    // the step's window: D1 rows read +-2 points inside, the closure rows the 6 points
    // next to a side
    const int wsx = min(sx - kAcSHL, nx - 8), wsz = sz - kAcH;
    const int s0 = grp * a.spb, s1 = min(s0 + a.spb, a.nshots);
    ac_window_x(a.curr + (long long)s0 * a.shot_stride, sA[0], wsx, wsz, nx, ldz);
    ac_window_x(a.f0 + (long long)s0 * a.fshot_stride, sF[0], wsx, wsz, nx, ldz);
    const int q = threadIdx.x;
    const int gx = sx + q / (kAcTZ / 4), z0 = sz + 4 * (q % (kAcTZ / 4));
    const bool active = gx < ex && z0 < ez;
    const long long g = active ? (long long)gx * ldz + z0 : 0;
    float K[4], gk[4], gr[4];
    unpack(ld4(a.K + g), K);
    float* G = a.grad + (long long)grp * a.gslab_stride + g;
    unpack(ld4(G), gk);
    unpack(ld4(G + a.shot_stride), gr);
    float kw[4];  // the K term's weight w / (dtg K^2)
#pragma unroll
    for (int k = 0; k < 4; k++) kw[k] = active ? a.wk / (K[k] * K[k]) : 0.f;
    const int col = (active ? gx - wsx : 2) * kAcWZ;
    const int zw = (active ? z0 : sz) - wsz;
    __syncthreads();
    for (int sh = s0; sh < s1; sh++) {
        const int b = (sh - s0) & 1;
        const long long fo = (long long)sh * a.fshot_stride;
        if (sh + 1 < s1) {
            ac_window_x(a.curr + (long long)(sh + 1) * a.shot_stride, sA[b ^ 1], wsx, wsz, nx, ldz);
            ac_window_x(a.f0 + fo + a.fshot_stride, sF[b ^ 1], wsx, wsz, nx, ldz);
        }
        if (active) {
            float fm[4] = {0.f, 0.f, 0.f, 0.f}, fp[4] = {0.f, 0.f, 0.f, 0.f};
            if (a.mode != 1) unpack(ld4(a.fm + fo + g), fm);
            if (a.mode != 2) unpack(ld4(a.fp + fo + g), fp);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int gz = z0 + k;
                if (gz >= nz) continue;
                const float* pa = &sA[b][col + zw + k];
                const float* pf = &sF[b][col + zw + k];
                const float ax = d1<false>(pa, kAcWZ, gx, nx) * a.rdx, az = d1<false>(pa, 1, gz, nz) * a.rdz;
                const float fx = d1<false>(pf, kAcWZ, gx, nx) * a.rdx, fz = d1<false>(pf, 1, gz, nz) * a.rdz;
                const float d2 =
                    a.mode == 0 ? fp[k] - 2.f * pf[0] + fm[k] : (a.mode == 1 ? fp[k] - pf[0] : -pf[0] + fm[k]);
                gk[k] -= kw[k] * pa[0] * d2;
                gr[k] += a.wr * (ax * fx + az * fz);
            }
        }
        __syncthreads();  // the next shot's windows are in; this shot's are read
    }
    if (active) {
        st4(G, make_float4(gk[0], gk[1], gk[2], gk[3]));
        st4(G + a.shot_stride, make_float4(gr[0], gr[1], gr[2], gr[3]));
    }
}

// point sources of the acoustic operator: per touched cell, the sum of its contributions
// in injector order, times dt^2 K (the leapfrog's factor), over the boundary divisor and
// times the taper, as the reference's injection into the right-hand side becomes after
// the step; init: the initial condition prev = 1/2 dt^2 K inject(trace sample 0)
__global__ void k_ainject(const InjCell* cells, int ncell, const int* ptr, const long long* off, const float* w,
                          const float* trace, int it, float* u, long long shot_stride, int ldz, const float* kdt2,
                          const float* div, const float* taper_x, const float* taper_z, int init)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const InjCell cell = cells[c];
    float f = 0.f;
    for (int k = ptr[c]; k < ptr[c + 1]; k++) f += w[k] * trace[off[k] + it];
    const long long g = (long long)cell.gx * ldz + cell.gz;
    float* p = u + cell.shot * shot_stride + g;
    if (init) *p += 0.5f * kdt2[g] * f;
    else *p += ((kdt2[g] * f / div[g]) * taper_z[cell.gz]) * taper_x[cell.gx];
}

__global__ void k_aextract(const float* u, long long shot_stride, int ldz, const PointEntry* ents, int nrec,
                           float* trace, int it)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const PointEntry e = ents[r];
    const float* p = u + e.shot * shot_stride;
    float v = 0.f;
    for (int ix = 0; ix < 3; ix++)
        for (int iz = 0; iz < 3; iz++) v += e.wx[ix] * e.wz[iz] * p[(long long)(e.xi0 + ix) * ldz + e.zi0 + iz];
    trace[e.trace + it] = v;
}

void launch_astep(const AcousticArgs& a, hipStream_t s)
{
    const int groups = (a.nshots + a.spb - 1) / a.spb;
    const int nin = (a.itx1 - a.itx0) * (a.itz1 - a.itz0), ne = (a.ntx * a.ntz - nin) * groups;
    const dim3 grid(((ne + 7) & ~7) + nin * groups);
    if (a.img) hipLaunchKernelGGL(k_astep<true>, grid, dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL(k_astep<false>, grid, dim3(kThreads), 0, s, a);
}

void launch_agrad(const AcousticArgs& a, hipStream_t s)
{
    const int groups = (a.nshots + a.spb - 1) / a.spb;
    hipLaunchKernelGGL(k_agrad, dim3(a.ntx * a.ntz * groups), dim3(kThreads), 0, s, a);
}

void launch_ainject(const InjCell* cells, int ncell, const int* ptr, const long long* off, const float* w,
                    const float* trace, int it, float* u, long long shot_stride, int ldz, const float* kdt2,
                    const float* div, const float* taper_x, const float* taper_z, bool init, hipStream_t s)
{
    if (ncell <= 0) return;
    hipLaunchKernelGGL(k_ainject, dim3((ncell + 127) / 128), dim3(128), 0, s, cells, ncell, ptr, off, w, trace, it, u,
                       shot_stride, ldz, kdt2, div, taper_x, taper_z, init ? 1 : 0);
}

void launch_aextract(const float* u, long long shot_stride, int ldz, const PointEntry* ents, int nrec, float* trace,
                     int it, hipStream_t s)
{
    if (nrec <= 0) return;
    hipLaunchKernelGGL(k_aextract, dim3((nrec + 127) / 128), dim3(128), 0, s, u, shot_stride, ldz, ents, nrec, trace,
                       it);
}

}  // namespace fwi2d
