// Device kernels of the elastic SBP-SAT propagation engine, written for gfx950.
//
// Layout: every field is [x][z] with z contiguous (index ix*nz + iz), as in the
// reference.  A workgroup of 256 threads owns a tile of 64 (z) x TX (x) cells for one
// shot; lanes run along z so every global and LDS access of a wavefront is a
// contiguous 64-float row.
//
// Tiling: tiles start at multiples of the tile size except the last one in each
// direction, which is anchored to end at the grid edge (start = n - T).  Each tile
// writes the cells in [start, next tile start) (the last tile writes to the edge), so
// writes never overlap, the SBP closure rows near the far edge are always written by
// the anchored tile, and a halo of 4 cells holds every stencil input.
#include "kernels.h"
#include "sbp_tables.h"

namespace fwi2d {

__constant__ float cD1B[4][6];
__constant__ float cD1I[2];
__constant__ float cSC[4];
__constant__ float cD2I[10];
__constant__ float cD2B[6][8][8];

void upload_constants(hipStream_t s)
{
    float d1b[4][6], d1i[2], sc[4], d2i[10], d2b[6][8][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 6; j++) d1b[i][j] = (float)kD1Closure[i][j];
    d1i[0] = (float)kD1Interior[0];
    d1i[1] = (float)kD1Interior[1];
    for (int i = 0; i < 4; i++) sc[i] = (float)kBoundaryDeriv[i];
    for (int i = 0; i < 10; i++) d2i[i] = (float)kD2Interior[i];
    for (int r = 0; r < 6; r++)
        for (int j = 0; j < 8; j++)
            for (int k = 0; k < 8; k++) d2b[r][j][k] = 0.f;
    for (const auto& e : kD2ClosureNZ) d2b[e.row][e.point][e.coef] = (float)(e.num / e.den);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD1B), d1b, sizeof(d1b), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD1I), d1i, sizeof(d1i), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cSC), sc, sizeof(sc), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD2I), d2i, sizeof(d2i), 0, hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(cD2B), d2b, sizeof(d2b), 0, hipMemcpyHostToDevice, s);
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
__device__ __forceinline__ void product_range(int s, int e, int n, int& lo, int& hi)
{
    lo = max(0, s - 2);
    hi = min(n, e + 2);
    if (s < 4) { lo = 0; hi = max(hi, 6); }
    if (e > n - 4) { lo = min(lo, n - 6); hi = n; }
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous run
// of (shot, tile) work items so that neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = bid & 7, local = bid >> 3;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
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

__device__ __noinline__ float d1_closure(const float* p, int s, int g, int n)
{
    float v = 0.f;
    if (g < 4) {
        const float* q = p - g * s;
        for (int j = 0; j < 6; j++) v += cD1B[g][j] * q[j * s];
    } else {
        const int r = n - 1 - g;
        const float* q = p + r * s;
        for (int j = 0; j < 6; j++) v -= cD1B[r][j] * q[-j * s];
    }
    return v;
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

__device__ __forceinline__ float d2_apply(const W5& w, const float* pu, int s)
{
    return w.c * pu[0] + w.m1 * pu[-s] + w.p1 * pu[s] + w.m2 * pu[-2 * s] + w.p2 * pu[2 * s];
}

// closure rows of D2(c) u (the 6x8x8 table)
__device__ __noinline__ float d2_closure(const float* pu, const float* pc, int s, int g, int n)
{
    float out = 0.f;
    const bool top = g < 6;
    const int r = top ? g : n - 1 - g;
    const int dir = top ? s : -s;
    const float* qu = top ? pu - g * s : pu + r * s;
    const float* qc = top ? pc - g * s : pc + r * s;
    for (int j = 0; j < 8; j++) {
        float val = 0.f;
        for (int k = 0; k < 8; k++) val += cD2B[r][j][k] * qc[k * dir];
        out += val * qu[j * dir];
    }
    return out;
}

// one-sided outward boundary derivative sum_k SC[k] u[k*s] (times h)
__device__ __forceinline__ float bnd_deriv(const float* p, int s)
{
    return cSC[0] * p[0] + cSC[1] * p[s] + cSC[2] * p[2 * s] + cSC[3] * p[3 * s];
}

// sum_j B[0][j] u[j*s]: first closure row (times h), used by the normal complement terms
__device__ __forceinline__ float d1_row0(const float* p, int s)
{
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < 6; j++) v += cD1B[0][j] * p[j * s];
    return v;
}

// ---------------------------------------------------------------------------
// Fused time step
//
// A workgroup owns a 64 (z) x TX (x) tile of one shot; each thread owns quads of 4
// consecutive z cells.  Phase 1 stages the state and mu with a 4-cell halo in LDS,
// phase 2 forms the products of first derivatives, phase 3 computes the new level.
// Quads whose stencils are all interior take the vectorised path; quads touching
// closure rows or SAT lines take the per-cell path.
// ---------------------------------------------------------------------------
template <int TX>
struct StepShape {
    static constexpr int TZ = kTileZ, H = 4;
    static constexpr int WZ = TZ + 2 * H, WX = TX + 2 * H;  // state/coefficient window (cells)
    static constexpr int WQ = WZ / 4;                        // window quads per column
    static constexpr int IX = TX + 4;                        // product columns (z extent = WZ)
    static constexpr int NQ1 = (WX * WQ + kThreads - 1) / kThreads;  // window quads per thread
    static constexpr int NQ2 = (IX * WQ + kThreads - 1) / kThreads;  // product quads per thread
    static constexpr int NQ3 = (TX * 16) / kThreads;                  // new-level quads per thread
};

struct TileInfo {
    int shot, tile, sx, ex, sz, ez, wsx, wsz, plx, phx, plz, phz;
};

struct StepSmem {
    float *ux, *uz, *mu, *mux, *px, *pz, *qx, *qz;
};

// ---- per-cell path (closure rows, SAT lines) ------------------------------------
// products at one cell; w = window index, returns (px, pz, qx, qz)
template <bool VTI, int TX>
__device__ __forceinline__ float4 product_cell(const StepArgs& a, const StepSmem& sm, int w, int gx, int gz,
                                               float lam, float c13, float eps)
{
    using S = StepShape<TX>;
    const float dxux = d1<false>(&sm.ux[w], S::WZ, gx, a.nx) * a.rdx;
    const float dzuz = d1<false>(&sm.uz[w], 1, gz, a.nz) * a.rdz;
    const float dxuz = d1<false>(&sm.uz[w], S::WZ, gx, a.nx) * a.rdx;
    const float dzux = d1<false>(&sm.ux[w], 1, gz, a.nz) * a.rdz;
    const float mu = sm.mu[w];
    float px, pz = 0.f;
    if (!VTI) {
        px = lam * (dxux + dzuz);
    } else {
        const float lamx = (1.f + 2.f * eps) * lam;
        px = lamx * dxux + c13 * dzuz;
        pz = lam * dzuz + c13 * dxux;
    }
    return make_float4(px, pz, mu * dxuz, mu * dzux);
}

// right-hand side r = (rx, rz) of one cell before sources: volume terms (we_op.cpp:690-720,
// 1665-1700) and SAT terms (we_op.cpp:726-822, 1705-1800)
template <bool VTI, int TX>
__device__ __forceinline__ void rhs_cell(const StepArgs& a, const StepSmem& sm, const TileInfo& t, int gx, int gz,
                                         float prx, float prz, float& rxo, float& rzo)
{
    using S = StepShape<TX>;
    const int nx = a.nx, nz = a.nz;
    const int w = (gx - t.wsx) * S::WZ + (gz - t.wsz);
    const int q = (gx - t.plx) * S::WZ + (gz - t.wsz);
    float rx, rz, d2zx, d2xz;
    if (gx >= 6 && gx < nx - 6) {
        const W5 wx = d2_weights(&sm.mu[w], S::WZ);
        rx = 2.f * d2_apply(VTI ? d2_weights(&sm.mux[w], S::WZ) : wx, &sm.ux[w], S::WZ);
        d2xz = d2_apply(wx, &sm.uz[w], S::WZ);
    } else {
        rx = 2.f * d2_closure(&sm.ux[w], VTI ? &sm.mux[w] : &sm.mu[w], S::WZ, gx, nx);
        d2xz = d2_closure(&sm.uz[w], &sm.mu[w], S::WZ, gx, nx);
    }
    if (gz >= 6 && gz < nz - 6) {
        const W5 wz = d2_weights(&sm.mu[w], 1);
        rz = 2.f * d2_apply(wz, &sm.uz[w], 1);
        d2zx = d2_apply(wz, &sm.ux[w], 1);
    } else {
        rz = 2.f * d2_closure(&sm.uz[w], &sm.mu[w], 1, gz, nz);
        d2zx = d2_closure(&sm.ux[w], &sm.mu[w], 1, gz, nz);
    }
    rx *= a.rdx2;
    rz *= a.rdz2;
    rz += d2xz * a.rdx2;
    rx += d2zx * a.rdz2;
    rx += d1<false>(&sm.px[q], S::WZ, gx, nx) * a.rdx;
    rz += d1<false>(VTI ? &sm.pz[q] : &sm.px[q], 1, gz, nz) * a.rdz;
    rz += d1<false>(&sm.qz[q], S::WZ, gx, nx) * a.rdx;
    rx += d1<false>(&sm.qx[q], 1, gz, nz) * a.rdz;

    const bool edge = (gz == 0) | (gz == nz - 1) | (gx == 0) | (gx == nx - 1);
    if (edge) {
        const int g = gx * a.ldz + gz;
        const float mu0 = sm.mu[w];
        const float lam0 = a.lam[g];
        const float rho0 = a.rho[g];
        const float f1tb = VTI ? a.c13[g] : lam0;
        const float epsf = VTI ? 1.f + 2.f * a.eps[g] : 1.f;
        for (int side = 0; side < 2; side++) {
            const int bc = side ? a.bc_bottom : a.bc_top;
            if (bc == 0 || gz != (side ? nz - 1 : 0)) continue;
            const float sgn = side ? 1.f : -1.f;
            const int sn = side ? -1 : 1;  // inward step along z
            float tt = sgn * mu0 * d1<false>(&sm.uz[w], S::WZ, gx, nx) * a.rdx + mu0 * bnd_deriv(&sm.ux[w], sn) * a.rdz;
            if (bc == 2) tt -= (0.5f * sqrtf(mu0 * rho0)) * prx / a.dt;
            rx -= a.sat_z * tt;
            tt = sgn * f1tb * d1<false>(&sm.ux[w], S::WZ, gx, nx) * a.rdx + (2.f * mu0) * bnd_deriv(&sm.uz[w], sn) * a.rdz;
            if (bc == 2) tt -= (0.5f * sqrtf((lam0 + 2.f * mu0) * rho0)) * prz / a.dt;
            rz -= a.sat_z * tt;
            rz += (a.sat_z * a.rdz) * lam0 * d1_row0(&sm.uz[w], sn);
        }
        for (int side = 0; side < 2; side++) {
            const int bc = side ? a.bc_right : a.bc_left;
            if (bc == 0 || gx != (side ? nx - 1 : 0)) continue;
            const float sgn = side ? 1.f : -1.f;
            const int sn = side ? -S::WZ : S::WZ;  // inward step along x
            const float f2 = epsf * (2.f * mu0);
            float tt = sgn * f1tb * d1<false>(&sm.uz[w], 1, gz, nz) * a.rdz + f2 * bnd_deriv(&sm.ux[w], sn) * a.rdx;
            if (bc == 2) tt -= (0.5f * sqrtf(epsf * (lam0 + 2.f * mu0) * rho0)) * prx / a.dt;
            rx -= a.sat_x * tt;
            rx += (a.sat_x * a.rdx) * (epsf * lam0) * d1_row0(&sm.ux[w], sn);
            tt = sgn * mu0 * d1<false>(&sm.ux[w], 1, gz, nz) * a.rdz + mu0 * bnd_deriv(&sm.uz[w], sn) * a.rdx;
            if (bc == 2) tt -= (0.5f * sqrtf(mu0 * rho0)) * prz / a.dt;
            rz -= a.sat_x * tt;
        }
    }
    rxo = rx;
    rzo = rz;
}

// ---- vectorised path: interior quad ------------------------------------------------
// D1 of a 12-cell column (cells z0-4 .. z0+7) at the 4 quad cells
__device__ __forceinline__ void d1_quad_z(const float* c, float* out)
{
#pragma unroll
    for (int k = 0; k < 4; k++)
        out[k] = cD1I[0] * (c[5 + k] - c[3 + k]) + cD1I[1] * (c[6 + k] - c[2 + k]);
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

__device__ __forceinline__ void load_col12(const float* base, int i, float* c)
{
    unpack(ld4(base + i - 4), c);
    unpack(ld4(base + i), c + 4);
    unpack(ld4(base + i + 4), c + 8);
}

template <bool VTI, int TX>
__device__ __forceinline__ void rhs_quad(const StepArgs& a, const StepSmem& sm, const TileInfo& t, int gx, int z0,
                                         float* rx, float* rz)
{
    using S = StepShape<TX>;
    const int w = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
    const int q = (gx - t.plx) * S::WZ + (z0 - t.wsz);
    // D2 along x of ux (coefficient 2mu, or (1+2eps)2mu) and uz (mu)
    {
        float U[5][4], Z[5][4], M[5][4], MX[VTI ? 5 : 1][4];
#pragma unroll
        for (int c = 0; c < 5; c++) {
            unpack(ld4(&sm.ux[w + (c - 2) * S::WZ]), U[c]);
            unpack(ld4(&sm.uz[w + (c - 2) * S::WZ]), Z[c]);
            unpack(ld4(&sm.mu[w + (c - 2) * S::WZ]), M[c]);
            if (VTI) unpack(ld4(&sm.mux[w + (c - 2) * S::WZ]), MX[c]);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float mk[5] = {M[0][k], M[1][k], M[2][k], M[3][k], M[4][k]};
            const float uk[5] = {U[0][k], U[1][k], U[2][k], U[3][k], U[4][k]};
            const float zk[5] = {Z[0][k], Z[1][k], Z[2][k], Z[3][k], Z[4][k]};
            const W5 wx = d2_weights(&mk[2], 1);
            if (VTI) {
                const float xk[5] = {MX[0][k], MX[1][k], MX[2][k], MX[3][k], MX[4][k]};
                rx[k] = 2.f * d2_apply(d2_weights(&xk[2], 1), &uk[2], 1);
            } else {
                rx[k] = 2.f * d2_apply(wx, &uk[2], 1);
            }
            rz[k] = d2_apply(wx, &zk[2], 1);  // D2x[mu] uz, added after D2z below
        }
    }
    // D2 along z
    float d2zx[4], d2zz[4];
    {
        float Uc[12], Zc[12], Mc[12];
        load_col12(sm.ux, w, Uc);
        load_col12(sm.uz, w, Zc);
        load_col12(sm.mu, w, Mc);
        d2_quad_z(Mc, Zc, Uc, d2zz, d2zx);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float d2xz = rz[k];
        rx[k] *= a.rdx2;
        rz[k] = 2.f * d2zz[k] * a.rdz2;
        rz[k] += d2xz * a.rdx2;
        rx[k] += d2zx[k] * a.rdz2;
    }
    // first derivatives of the products
    {
        float P[4][4], Q[4][4];
        const int off[4] = {-2, -1, 1, 2};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            unpack(ld4(&sm.px[q + off[c] * S::WZ]), P[c]);
            unpack(ld4(&sm.qz[q + off[c] * S::WZ]), Q[c]);
        }
        float Pc[12], Qc[12], dpz[4], dqz[4];
        load_col12(VTI ? sm.pz : sm.px, q, Pc);
        load_col12(sm.qx, q, Qc);
        d1_quad_z(Pc, dpz);
        d1_quad_z(Qc, dqz);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float dpx = cD1I[0] * (P[2][k] - P[1][k]) + cD1I[1] * (P[3][k] - P[0][k]);
            const float dqx = cD1I[0] * (Q[2][k] - Q[1][k]) + cD1I[1] * (Q[3][k] - Q[0][k]);
            rx[k] += dpx * a.rdx;
            rz[k] += dpz[k] * a.rdz;
            rz[k] += dqx * a.rdx;
            rx[k] += dqz[k] * a.rdz;
        }
    }
}

// products of one interior quad (all four cells have interior D1 stencils)
template <bool VTI, int TX>
__device__ __forceinline__ void product_quad(const StepArgs& a, const StepSmem& sm, int w, const float* lam,
                                             const float* c13, const float* eps, float4& P, float4& PZ, float4& QX,
                                             float4& QZ)
{
    using S = StepShape<TX>;
    float X[4][4], XZ[4][4];
    const int off[4] = {-2, -1, 1, 2};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        unpack(ld4(&sm.ux[w + off[c] * S::WZ]), X[c]);
        unpack(ld4(&sm.uz[w + off[c] * S::WZ]), XZ[c]);
    }
    float Uc[12], Zc[12], dzux[4], dzuz[4], mu[4];
    load_col12(sm.ux, w, Uc);
    load_col12(sm.uz, w, Zc);
    d1_quad_z(Uc, dzux);
    d1_quad_z(Zc, dzuz);
    unpack(ld4(&sm.mu[w]), mu);
    float p[4], pz[4], qx[4], qz[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float dxux = (cD1I[0] * (X[2][k] - X[1][k]) + cD1I[1] * (X[3][k] - X[0][k])) * a.rdx;
        const float dxuz = (cD1I[0] * (XZ[2][k] - XZ[1][k]) + cD1I[1] * (XZ[3][k] - XZ[0][k])) * a.rdx;
        const float duz = dzuz[k] * a.rdz;
        const float dux = dzux[k] * a.rdz;
        if (!VTI) {
            p[k] = lam[k] * (dxux + duz);
            pz[k] = 0.f;
        } else {
            const float lamx = (1.f + 2.f * eps[k]) * lam[k];
            p[k] = lamx * dxux + c13[k] * duz;
            pz[k] = lam[k] * duz + c13[k] * dxux;
        }
        qx[k] = mu[k] * dxuz;
        qz[k] = mu[k] * dux;
    }
    P = make_float4(p[0], p[1], p[2], p[3]);
    PZ = make_float4(pz[0], pz[1], pz[2], pz[3]);
    QX = make_float4(qx[0], qx[1], qx[2], qx[3]);
    QZ = make_float4(qz[0], qz[1], qz[2], qz[3]);
}

template <bool VTI, int TX>
__global__ void __launch_bounds__(kThreads) k_step(const StepArgs a)
{
    using S = StepShape<TX>;
    __shared__ __attribute__((aligned(16))) float sUx[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sUz[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sMu[S::WX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sMuX[VTI ? S::WX * S::WZ : 4];  // (1+2eps) mu
    __shared__ __attribute__((aligned(16))) float sPx[S::IX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sPz[VTI ? S::IX * S::WZ : 4];
    __shared__ __attribute__((aligned(16))) float sQx[S::IX * S::WZ];
    __shared__ __attribute__((aligned(16))) float sQz[S::IX * S::WZ];
    const StepSmem sm{sUx, sUz, sMu, sMuX, sPx, sPz, sQx, sQz};

    const int ntiles = a.ntx * a.ntz;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    TileInfo t;
    t.shot = lin / ntiles;
    t.tile = lin - t.shot * ntiles;
    const int tix = t.tile / a.ntz;
    const int tiz = t.tile - tix * a.ntz;
    const int nx = a.nx, nz = a.nz, ldz = a.ldz;
    tile_range_x_hd(tix, a.ntx, TX, nx, t.sx, t.ex);
    tile_range_z_hd(tiz, a.ntz, nz, ldz, t.sz, t.ez);
    t.wsx = window_start(t.sx, S::H);
    t.wsz = window_start(t.sz, S::H);
    product_range(t.sx, t.ex, nx, t.plx, t.phx);
    product_range(t.sz, t.ez, nz, t.plz, t.phz);
    const long long nxz = (long long)nx * ldz;
    const float* cux = a.curr + t.shot * a.shot_stride;
    const float* cuz = cux + nxz;
    const float* pux = a.prev + t.shot * a.shot_stride;
    const int tid = threadIdx.x;

    // ---- phase 1: every global load of the thread is issued first ----
    float4 wux[S::NQ1], wuz[S::NQ1], wmu[S::NQ1], wep[VTI ? S::NQ1 : 1];
#pragma unroll
    for (int m = 0; m < S::NQ1; m++) {
        const int i = tid + m * kThreads;
        const int lx = i / S::WQ, lq = i - lx * S::WQ;
        const int gx = t.wsx + lx, gz = t.wsz + 4 * lq;
        const bool ok = i < S::WX * S::WQ && gx >= 0 && gx < nx && gz >= 0 && gz < ldz;
        const int g = ok ? gx * ldz + gz : 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        wux[m] = ok ? ld4(cux + g) : z4;
        wuz[m] = ok ? ld4(cuz + g) : z4;
        wmu[m] = ok ? ld4(a.mu + g) : z4;
        if (VTI) wep[m] = ok ? ld4(a.eps + g) : z4;
    }
    // products' coefficients
    float4 plam[S::NQ2], pc13[VTI ? S::NQ2 : 1], peps[VTI ? S::NQ2 : 1];
#pragma unroll
    for (int m = 0; m < S::NQ2; m++) {
        const int i = tid + m * kThreads;
        const int lc = i / S::WQ, lq = i - lc * S::WQ;
        const int gx = t.plx + lc, gz = t.wsz + 4 * lq;
        const bool ok = i < S::IX * S::WQ && gx < t.phx && gz + 3 >= t.plz && gz < t.phz;
        const int g = ok ? gx * ldz + gz : 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        plam[m] = ok ? ld4(a.lam + g) : z4;
        if (VTI) {
            pc13[m] = ok ? ld4(a.c13 + g) : z4;
            peps[m] = ok ? ld4(a.eps + g) : z4;
        }
    }
    // new-level cells: previous level, density, tapers
    float4 qprx[S::NQ3], qprz[S::NQ3], qrho[S::NQ3], qtz[S::NQ3];
    float qtx[S::NQ3];
#pragma unroll
    for (int m = 0; m < S::NQ3; m++) {
        const int i = tid + m * kThreads;
        const int gx = t.sx + (i >> 4), gz = t.sz + 4 * (i & 15);
        const bool ok = gx < t.ex && gz < t.ez;
        const int g = ok ? gx * ldz + gz : 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        qprx[m] = ok ? ld4(pux + g) : z4;
        qprz[m] = ok ? ld4(pux + nxz + g) : z4;
        qrho[m] = ok ? ld4(a.rho + g) : make_float4(1.f, 1.f, 1.f, 1.f);
        qtz[m] = ok ? ld4(a.taper_z + gz) : z4;
        qtx[m] = ok ? a.taper_x[gx] : 0.f;
    }
#pragma unroll
    for (int m = 0; m < S::NQ1; m++) {
        const int i = tid + m * kThreads;
        if (i < S::WX * S::WQ) {
            const int lx = i / S::WQ, lq = i - lx * S::WQ;
            const int o = lx * S::WZ + 4 * lq;
            st4(&sUx[o], wux[m]);
            st4(&sUz[o], wuz[m]);
            st4(&sMu[o], wmu[m]);
            if (VTI) {
                const float4 e = wep[m], mu = wmu[m];
                st4(&sMuX[o], make_float4((1.f + 2.f * e.x) * mu.x, (1.f + 2.f * e.y) * mu.y, (1.f + 2.f * e.z) * mu.z,
                                          (1.f + 2.f * e.w) * mu.w));
            }
        }
    }
    __syncthreads();

    // ---- extraction of the current level by the owning tile ----
    if (a.ext_ptr) {
        const int b = a.ext_ptr[t.shot * ntiles + t.tile], e = a.ext_ptr[t.shot * ntiles + t.tile + 1];
        for (int k = b * 2 + tid; k < e * 2; k += kThreads) {
            const PointEntry en = a.ext[a.ext_idx[k >> 1]];
            const int comp = k & 1;
            const float* su = comp ? sUz : sUx;
            float v = 0.f;
            for (int ix = 0; ix < 3; ix++)
                for (int iz = 0; iz < 3; iz++)
                    v += en.wx[ix] * en.wz[iz] * su[(en.xi0 + ix - t.wsx) * S::WZ + (en.zi0 + iz - t.wsz)];
            a.ext_trace[en.trace + comp * a.ext_comp_stride + a.it] = v;
        }
    }

    // ---- phase 2: products of first derivatives on [plx, phx) x [plz, phz) ----
#pragma unroll
    for (int m = 0; m < S::NQ2; m++) {
        const int i = tid + m * kThreads;
        if (i >= S::IX * S::WQ) break;
        const int lc = i / S::WQ, lq = i - lc * S::WQ;
        const int gx = t.plx + lc, z0 = t.wsz + 4 * lq;
        const int o = lc * S::WZ + 4 * lq;
        const int w = (gx - t.wsx) * S::WZ + 4 * lq;
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f), PZ = P, QX = P, QZ = P;
        if (gx < t.phx && z0 + 3 >= t.plz && z0 < t.phz) {
            const bool fast = gx >= 4 && gx < nx - 4 && z0 >= 4 && z0 + 3 < nz - 4 && lq >= 1 && lq + 1 < S::WQ &&
                              z0 >= t.plz && z0 + 3 < t.phz;
            float lam[4], c13[4] = {0.f, 0.f, 0.f, 0.f}, eps[4] = {0.f, 0.f, 0.f, 0.f};
            unpack(plam[m], lam);
            if (VTI) {
                unpack(pc13[m], c13);
                unpack(peps[m], eps);
            }
            if (fast) {
                product_quad<VTI, TX>(a, sm, w, lam, c13, eps, P, PZ, QX, QZ);
            } else {
                float p[4] = {0.f, 0.f, 0.f, 0.f}, pz[4] = {0.f, 0.f, 0.f, 0.f}, qx[4] = {0.f, 0.f, 0.f, 0.f},
                      qz[4] = {0.f, 0.f, 0.f, 0.f};
                for (int k = 0; k < 4; k++) {
                    const int gz = z0 + k;
                    if (gz < t.plz || gz >= t.phz) continue;
                    const float4 r = product_cell<VTI, TX>(a, sm, w + k, gx, gz, lam[k], c13[k], eps[k]);
                    p[k] = r.x;
                    pz[k] = r.y;
                    qx[k] = r.z;
                    qz[k] = r.w;
                }
                P = make_float4(p[0], p[1], p[2], p[3]);
                PZ = make_float4(pz[0], pz[1], pz[2], pz[3]);
                QX = make_float4(qx[0], qx[1], qx[2], qx[3]);
                QZ = make_float4(qz[0], qz[1], qz[2], qz[3]);
            }
        }
        st4(&sPx[o], P);
        if (VTI) st4(&sPz[o], PZ);
        st4(&sQx[o], QX);
        st4(&sQz[o], QZ);
    }
    __syncthreads();

    // ---- phase 3: new level ----
    float* nux = a.next + t.shot * a.shot_stride;
    float* nuz = nux + nxz;
    const int ib = a.inj_ptr[t.shot * ntiles + t.tile], ie = a.inj_ptr[t.shot * ntiles + t.tile + 1];
#pragma unroll
    for (int m = 0; m < S::NQ3; m++) {
        const int i = tid + m * kThreads;
        const int gx = t.sx + (i >> 4), z0 = t.sz + 4 * (i & 15);
        if (gx >= t.ex || z0 >= t.ez) continue;
        float prx[4], prz[4], rho[4], tz[4];
        unpack(qprx[m], prx);
        unpack(qprz[m], prz);
        unpack(qrho[m], rho);
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
        const bool fast = gx >= 6 && gx < nx - 6 && z0 >= 6 && z0 + 3 < nz - 6;
        if (fast) {
            rhs_quad<VTI, TX>(a, sm, t, gx, z0, rx, rz);
        } else {
            for (int k = 0; k < 4; k++) {
                rx[k] = rz[k] = 0.f;
                if (z0 + k < t.ez) rhs_cell<VTI, TX>(a, sm, t, gx, z0 + k, prx[k], prz[k], rx[k], rz[k]);
            }
        }
        // sources (we_op.cpp:722-724), added after the volume terms and before the SAT
        // terms in the reference; the SAT lines never carry sources' footprints away
        // from their cells, so the order only changes rounding
        for (int k2 = ib; k2 < ie; k2++) {
            const PointEntry& en = a.inj[a.inj_idx[k2]];
            const int ax = gx - en.xi0;
            if ((unsigned)ax >= 3u) continue;
            const float sx = a.inj_trace[en.trace + a.it];
            const float sz = a.inj_trace[en.trace + a.inj_comp_stride + a.it];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int az = z0 + k - en.zi0;
                if ((unsigned)az < 3u) {
                    const float wgt = en.wx[ax] * en.wz[az];
                    rx[k] += wgt * sx;
                    rz[k] += wgt * sz;
                }
            }
        }
        float vx[4], vz[4];
        const int wq = (gx - t.wsx) * S::WZ + (z0 - t.wsz);
        float cx[4], cz[4];
        unpack(ld4(&sUx[wq]), cx);
        unpack(ld4(&sUz[wq]), cz);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            // leapfrog (we_op.cpp:824-830)
            vx[k] = a.dt2 * rx[k] / rho[k] + 2.f * cx[k] - prx[k];
            vz[k] = a.dt2 * rz[k] / rho[k] + 2.f * cz[k] - prz[k];
        }
        if (!fast) {
            // implicit part of the absorbing SAT on the perimeter (spatial_operators.cpp:307-521)
            for (int k = 0; k < 4; k++) {
                const int gz = z0 + k;
                if (gz >= t.ez) continue;
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
    }
}

// ---------------------------------------------------------------------------
// Initial condition: prev = 1/2 dt^2/rho * inject(trace[0])  (we_op.cpp:655-663)
// ---------------------------------------------------------------------------
__global__ void k_init_prev(float* prev, long long shot_stride, long long nxz, int ldz, const int* shot_ptr,
                            const PointEntry* ents, const float* trace, long long comp_stride)
{
    // one single-thread block per shot: deterministic accumulation order over its injectors
    const int s = blockIdx.x;
    const int b = shot_ptr[s], e = shot_ptr[s + 1];
    float* px = prev + s * shot_stride;
    float* pz = px + nxz;
    for (int k = b; k < e; k++) {
        const PointEntry en = ents[k];
        const float vx = trace[en.trace], vz = trace[en.trace + comp_stride];
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
                          int n, float* trace, long long comp_stride, int it)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * n) return;
    const PointEntry en = ents[k >> 1];
    const int comp = k & 1;
    const float* u = state + en.shot * shot_stride + comp * nxz;
    float v = 0.f;
    for (int ix = 0; ix < 3; ix++)
        for (int iz = 0; iz < 3; iz++) v += en.wx[ix] * en.wz[iz] * u[(long long)(en.xi0 + ix) * ldz + en.zi0 + iz];
    trace[en.trace + comp * comp_stride + it] = v;
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

template <bool VTI>
static void step_dispatch(const StepArgs& a, int tile_x, hipStream_t s)
{
    const int nb = a.ntx * a.ntz * a.nshots;
    if (tile_x == 32) hipLaunchKernelGGL((k_step<VTI, 32>), dim3(nb), dim3(kThreads), 0, s, a);
    else hipLaunchKernelGGL((k_step<VTI, 16>), dim3(nb), dim3(kThreads), 0, s, a);
}

void launch_step(const StepArgs& a, bool vti, int tile_x, hipStream_t s)
{
    if (vti) step_dispatch<true>(a, tile_x, s);
    else step_dispatch<false>(a, tile_x, s);
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
                      const PointEntry* ents, const float* trace, long long comp_stride, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_prev, dim3(nshots), dim3(1), 0, s, prev, shot_stride, nxz, ldz, shot_ptr, ents, trace,
                       comp_stride);
}

void launch_scale_prev(float* prev, const float* rho, float half_dt2, long long nxz, int nshots, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_prev, dim3(grid_for(2 * nxz * nshots)), dim3(kThreads), 0, s, prev, rho, half_dt2, nxz,
                       nshots);
}

void launch_extract(const float* state, long long shot_stride, long long nxz, int ldz, const PointEntry* ents, int n,
                    float* trace, long long comp_stride, int it, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_extract, dim3((2 * n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, state, shot_stride,
                       nxz, ldz, ents, n, trace, comp_stride, it);
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

}  // namespace fwi2d
