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

int tiles_along(int n, int t) { return (n + t - 1) / t; }

// ---------------------------------------------------------------------------
// Tile geometry
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_range(int k, int ntile, int T, int n, int& s, int& e)
{
    if (ntile == 1) { s = 0; e = n; return; }
    if (k == ntile - 1) { s = n - T; e = n; return; }
    s = k * T;
    e = min(s + T, n - T);
}

// first LDS row of a tile's window with halo H such that SBP closures at the near edge
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

// ---------------------------------------------------------------------------
// SBP stencils on an LDS line: p points at the element of global index g (0 <= g < n),
// s is the LDS stride along the differentiated axis.  Results are NOT divided by h.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float d1(const float* p, int s, int g, int n)
{
    if (g >= 4 && g < n - 4)
        return cD1I[0] * (p[s] - p[-s]) + cD1I[1] * (p[2 * s] - p[-2 * s]);
    float v = 0.f;
    if (g < 4) {
        const float* q = p - g * s;
#pragma unroll
        for (int j = 0; j < 6; j++) v += cD1B[g][j] * q[j * s];
    } else {
        const int r = n - 1 - g;
        const float* q = p + r * s;
#pragma unroll
        for (int j = 0; j < 6; j++) v -= cD1B[r][j] * q[-j * s];
    }
    return v;
}

// variable-coefficient second derivative D2(c) u at g (Dzz_var/Dxx_var rows)
__device__ __forceinline__ float d2(const float* pu, const float* pc, int s, int g, int n)
{
    if (g >= 6 && g < n - 6) {
        const float fm2 = pc[-2 * s], fm1 = pc[-s], f0 = pc[0], fp1 = pc[s], fp2 = pc[2 * s];
        return (cD2I[0] * f0 + cD2I[1] * (fm1 + fp1) + cD2I[2] * (fm2 + fp2)) * pu[0]
             + (cD2I[3] * fm2 + cD2I[4] * fm1 + cD2I[5] * f0 + cD2I[6] * fp1) * pu[-s]
             + (cD2I[3] * fp2 + cD2I[4] * fp1 + cD2I[5] * f0 + cD2I[6] * fm1) * pu[s]
             + (cD2I[7] * fm2 + cD2I[8] * fm1 + cD2I[9] * f0) * pu[-2 * s]
             + (cD2I[7] * fp2 + cD2I[8] * fp1 + cD2I[9] * f0) * pu[2 * s];
    }
    float out = 0.f;
    if (g < 6) {
        const float* qu = pu - g * s;
        const float* qc = pc - g * s;
        for (int j = 0; j < 8; j++) {
            float val = 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) val += cD2B[g][j][k] * qc[k * s];
            out += val * qu[j * s];
        }
    } else {
        const int r = n - 1 - g;
        const float* qu = pu + r * s;
        const float* qc = pc + r * s;
        for (int j = 0; j < 8; j++) {
            float val = 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) val += cD2B[r][j][k] * qc[-k * s];
            out += val * qu[-j * s];
        }
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
// ---------------------------------------------------------------------------
template <bool VTI, int TX>
__global__ void __launch_bounds__(kThreads) k_step(const StepArgs a)
{
    constexpr int TZ = kTileZ, H = 4;
    constexpr int WZ = TZ + 2 * H, WX = TX + 2 * H;  // state/coefficient window
    constexpr int IZ = TZ + 4, IX = TX + 4;          // first-derivative products window
    __shared__ float sUx[WX * WZ];
    __shared__ float sUz[WX * WZ];
    __shared__ float sMu[WX * WZ];
    __shared__ float sMuX[VTI ? WX * WZ : 1];        // (1+2eps) mu
    __shared__ float sPx[IX * IZ];
    __shared__ float sPz[VTI ? IX * IZ : 1];
    __shared__ float sQx[IX * IZ];
    __shared__ float sQz[IX * IZ];

    const int ntiles = a.ntx * a.ntz;
    const int lin = xcd_remap(blockIdx.x, gridDim.x);
    const int shot = lin / ntiles;
    const int tile = lin - shot * ntiles;
    const int tix = tile / a.ntz;
    const int tiz = tile - tix * a.ntz;
    int sx, ex, sz, ez;
    tile_range(tix, a.ntx, TX, a.nx, sx, ex);
    tile_range(tiz, a.ntz, TZ, a.nz, sz, ez);
    const int wsx = window_start(sx, H), wsz = window_start(sz, H);
    const int nx = a.nx, nz = a.nz;
    const long long nxz = (long long)nx * nz;
    const float* cux = a.curr + shot * a.shot_stride;
    const float* cuz = cux + nxz;

    // ---- phase 1: state and coefficient window ----
    for (int i = threadIdx.x; i < WX * WZ; i += kThreads) {
        const int lx = i / WZ, lz = i - lx * WZ;
        const int gx = wsx + lx, gz = wsz + lz;
        float vx = 0.f, vz = 0.f, m = 0.f, mx = 0.f;
        if (gx >= 0 && gx < nx && gz >= 0 && gz < nz) {
            const long long g = (long long)gx * nz + gz;
            vx = cux[g];
            vz = cuz[g];
            m = a.mu[g];
            if (VTI) mx = (1.f + 2.f * a.eps[g]) * m;
        }
        sUx[i] = vx;
        sUz[i] = vz;
        sMu[i] = m;
        if (VTI) sMuX[i] = mx;
    }
    __syncthreads();

    // ---- extraction of the current level by the owning tile ----
    if (a.ext_ptr) {
        const int b = a.ext_ptr[shot * ntiles + tile], e = a.ext_ptr[shot * ntiles + tile + 1];
        for (int k = b * 2 + threadIdx.x; k < e * 2; k += kThreads) {
            const PointEntry en = a.ext[a.ext_idx[k >> 1]];
            const int comp = k & 1;
            const float* su = comp ? sUz : sUx;
            float v = 0.f;
            for (int ix = 0; ix < 3; ix++)
                for (int iz = 0; iz < 3; iz++)
                    v += en.wx[ix] * en.wz[iz] * su[(en.xi0 + ix - wsx) * WZ + (en.zi0 + iz - wsz)];
            a.ext_trace[en.trace + comp * a.ext_comp_stride + a.it] = v;
        }
    }

    // ---- phase 2: products of first derivatives on the range the tile reads ----
    int plx, phx, plz, phz;
    product_range(sx, ex, nx, plx, phx);
    product_range(sz, ez, nz, plz, phz);
    for (int i = threadIdx.x; i < IX * IZ; i += kThreads) {
        const int lx = i / IZ, lz = i - lx * IZ;
        const int gx = plx + lx, gz = plz + lz;
        float px = 0.f, pz = 0.f, qx = 0.f, qz = 0.f;
        if (gx < phx && gz < phz) {
            const int w = (gx - wsx) * WZ + (gz - wsz);
            const long long g = (long long)gx * nz + gz;
            const float dxux = d1(&sUx[w], WZ, gx, nx) * a.rdx;
            const float dzuz = d1(&sUz[w], 1, gz, nz) * a.rdz;
            const float dxuz = d1(&sUz[w], WZ, gx, nx) * a.rdx;
            const float dzux = d1(&sUx[w], 1, gz, nz) * a.rdz;
            const float lam = a.lam[g];
            const float mu = sMu[w];
            if (!VTI) {
                px = lam * (dxux + dzuz);
            } else {
                const float c13 = a.c13[g];
                const float lamx = (1.f + 2.f * a.eps[g]) * lam;
                px = lamx * dxux + c13 * dzuz;
                pz = lam * dzuz + c13 * dxux;
            }
            qx = mu * dxuz;
            qz = mu * dzux;
        }
        sPx[i] = px;
        if (VTI) sPz[i] = pz;
        sQx[i] = qx;
        sQz[i] = qz;
    }
    __syncthreads();

    // ---- phase 3: new level on the tile ----
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int gz = sz + lane;
    if (gz >= ez) return;
    const float* pux = a.prev + shot * a.shot_stride;
    const float* puz = pux + nxz;
    float* nux = a.next + shot * a.shot_stride;
    float* nuz = nux + nxz;
    const int ib = a.inj_ptr[shot * ntiles + tile], ie = a.inj_ptr[shot * ntiles + tile + 1];
    const float tz = a.taper_z[gz];

#pragma unroll
    for (int j = 0; j < TX / 4; j++) {
        const int gx = sx + wave + 4 * j;
        if (gx >= ex) break;
        const int w = (gx - wsx) * WZ + (gz - wsz);
        const int q = (gx - plx) * IZ + (gz - plz);
        const long long g = (long long)gx * nz + gz;

        // volume terms, in the reference's order (we_op.cpp:690-720, 1665-1700)
        float rx = 2.f * d2(&sUx[w], VTI ? &sMuX[w] : &sMu[w], WZ, gx, nx) * a.rdx2;
        float rz = 2.f * d2(&sUz[w], &sMu[w], 1, gz, nz) * a.rdz2;
        rz += d2(&sUz[w], &sMu[w], WZ, gx, nx) * a.rdx2;
        rx += d2(&sUx[w], &sMu[w], 1, gz, nz) * a.rdz2;
        rx += d1(&sPx[q], IZ, gx, nx) * a.rdx;
        rz += d1(VTI ? &sPz[q] : &sPx[q], 1, gz, nz) * a.rdz;
        rz += d1(&sQz[q], IZ, gx, nx) * a.rdx;
        rx += d1(&sQx[q], 1, gz, nz) * a.rdz;

        // sources
        for (int k = ib; k < ie; k++) {
            const PointEntry& en = a.inj[a.inj_idx[k]];
            const int ax = gx - en.xi0, az = gz - en.zi0;
            if ((unsigned)ax < 3u && (unsigned)az < 3u) {
                const float wgt = en.wx[ax] * en.wz[az];
                rx += wgt * a.inj_trace[en.trace + a.it];
                rz += wgt * a.inj_trace[en.trace + a.inj_comp_stride + a.it];
            }
        }

        float prx = pux[g], prz = puz[g];
        if (a.taper_prev) {
            prx = (prx * tz) * a.taper_x[gx];
            prz = (prz * tz) * a.taper_x[gx];
        }

        // SAT boundary terms (we_op.cpp:726-822, 1705-1800)
        const bool edge = (gz == 0) | (gz == nz - 1) | (gx == 0) | (gx == nx - 1);
        if (edge) {
            const float mu0 = sMu[w];
            const float lam0 = a.lam[g];
            const float rho0 = a.rho[g];
            const float f1tb = VTI ? a.c13[g] : lam0;
            const float epsf = VTI ? 1.f + 2.f * a.eps[g] : 1.f;
            for (int side = 0; side < 2; side++) {
                const int bc = side ? a.bc_bottom : a.bc_top;
                if (bc == 0 || gz != (side ? nz - 1 : 0)) continue;
                const float sgn = side ? 1.f : -1.f;
                const int sn = side ? -1 : 1;  // inward step along z
                float t = sgn * mu0 * d1(&sUz[w], WZ, gx, nx) * a.rdx + mu0 * bnd_deriv(&sUx[w], sn) * a.rdz;
                if (bc == 2) t -= (0.5f * sqrtf(mu0 * rho0)) * prx / a.dt;
                rx -= a.sat_z * t;
                t = sgn * f1tb * d1(&sUx[w], WZ, gx, nx) * a.rdx + (2.f * mu0) * bnd_deriv(&sUz[w], sn) * a.rdz;
                if (bc == 2) t -= (0.5f * sqrtf((lam0 + 2.f * mu0) * rho0)) * prz / a.dt;
                rz -= a.sat_z * t;
                rz += (a.sat_z * a.rdz) * lam0 * d1_row0(&sUz[w], sn);
            }
            for (int side = 0; side < 2; side++) {
                const int bc = side ? a.bc_right : a.bc_left;
                if (bc == 0 || gx != (side ? nx - 1 : 0)) continue;
                const float sgn = side ? 1.f : -1.f;
                const int sn = side ? -WZ : WZ;  // inward step along x
                const float f2 = epsf * (2.f * mu0);
                float t = sgn * f1tb * d1(&sUz[w], 1, gz, nz) * a.rdz + f2 * bnd_deriv(&sUx[w], sn) * a.rdx;
                if (bc == 2) t -= (0.5f * sqrtf(epsf * (lam0 + 2.f * mu0) * rho0)) * prx / a.dt;
                rx -= a.sat_x * t;
                rx += (a.sat_x * a.rdx) * (epsf * lam0) * d1_row0(&sUx[w], sn);
                t = sgn * mu0 * d1(&sUx[w], 1, gz, nz) * a.rdz + mu0 * bnd_deriv(&sUz[w], sn) * a.rdx;
                if (bc == 2) t -= (0.5f * sqrtf(mu0 * rho0)) * prz / a.dt;
                rz -= a.sat_x * t;
            }
        }

        // leapfrog (we_op.cpp:824-830)
        const float rho = a.rho[g];
        float vx = a.dt2 * rx / rho + 2.f * sUx[w] - prx;
        float vz = a.dt2 * rz / rho + 2.f * sUz[w] - prz;

        // implicit part of the absorbing SAT on the perimeter (spatial_operators.cpp:307-521)
        if (edge) {
            float dvx = 1.f, dvz = 1.f;
            if (gz == 0) { dvx = a.div_tb[gx]; dvz = a.div_tb[nx + gx]; }
            else if (gz == nz - 1) { dvx = a.div_tb[2 * nx + gx]; dvz = a.div_tb[3 * nx + gx]; }
            else if (gx == 0) { dvx = a.div_lr[gz]; dvz = a.div_lr[nz + gz]; }
            else { dvx = a.div_lr[2 * nz + gz]; dvz = a.div_lr[3 * nz + gz]; }
            vx = vx / dvx;
            vz = vz / dvz;
        }
        // cos^2 taper of the new level (we_op.cpp:846-861)
        const float txf = a.taper_x[gx];
        nux[g] = (vx * tz) * txf;
        nuz[g] = (vz * tz) * txf;
    }
}

// ---------------------------------------------------------------------------
// Initial condition: prev = 1/2 dt^2/rho * inject(trace[0])  (we_op.cpp:655-663)
// ---------------------------------------------------------------------------
__global__ void k_init_prev(float* prev, long long shot_stride, int nz, const int* shot_ptr,
                            const PointEntry* ents, const float* trace, long long comp_stride, long long nxz)
{
    // one thread per shot: deterministic accumulation order over its injectors
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    (void)nxz;
    if (s >= (int)gridDim.x * (int)blockDim.x) return;
    const int b = shot_ptr[s], e = shot_ptr[s + 1];
    float* px = prev + s * shot_stride;
    float* pz = px + nxz;
    for (int k = b; k < e; k++) {
        const PointEntry en = ents[k];
        const float vx = trace[en.trace], vz = trace[en.trace + comp_stride];
        for (int ix = 0; ix < 3; ix++)
            for (int iz = 0; iz < 3; iz++) {
                const long long g = (long long)(en.xi0 + ix) * nz + (en.zi0 + iz);
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
__global__ void k_extract(const float* state, long long shot_stride, int nz, long long nxz, const PointEntry* ents,
                          int n, float* trace, long long comp_stride, int it)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * n) return;
    const PointEntry en = ents[k >> 1];
    const int comp = k & 1;
    const float* u = state + en.shot * shot_stride + comp * nxz;
    float v = 0.f;
    for (int ix = 0; ix < 3; ix++)
        for (int iz = 0; iz < 3; iz++) v += en.wx[ix] * en.wz[iz] * u[(long long)(en.xi0 + ix) * nz + en.zi0 + iz];
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
    int sx, ex, sz, ez;
    tile_range(tix, a.ntx, TX, a.nx, sx, ex);
    tile_range(tiz, a.ntz, TZ, a.nz, sz, ez);
    const int wsx = window_start(sx, H), wsz = window_start(sz, H);
    const int nx = a.nx, nz = a.nz;
    const long long nxz = (long long)nx * nz;
    const float* dA = (a.lag ? a.adj_prev : a.adj) + shot * a.shot_stride;
    const float* F0 = a.f0 + shot * a.fshot_stride;
    for (int i = threadIdx.x; i < WX * WZ; i += kThreads) {
        const int lx = i / WZ, lz = i - lx * WZ;
        const int gx = wsx + lx, gz = wsz + lz;
        float ax = 0.f, az = 0.f, fx = 0.f, fz = 0.f;
        if (gx >= 0 && gx < nx && gz >= 0 && gz < nz) {
            const long long g = (long long)gx * nz + gz;
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
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gz = sz + lane;
    if (gz >= ez) return;
    const float* A = a.adj + shot * a.shot_stride;
    const float* Fm = a.fm ? a.fm + shot * a.fshot_stride : nullptr;
    const float* Fp = a.fp ? a.fp + shot * a.fshot_stride : nullptr;
    float* G = a.grad + shot * a.gshot_stride;
#pragma unroll
    for (int j = 0; j < TX / 4; j++) {
        const int gx = sx + wave + 4 * j;
        if (gx >= ex) break;
        const int w = (gx - wsx) * WZ + (gz - wsz);
        const long long g = (long long)gx * nz + gz;
        const float axx = d1(&sAx[w], WZ, gx, nx) * a.rdx;  // adjoint x_x
        const float azx = d1(&sAx[w], 1, gz, nz) * a.rdz;   // adjoint x_z
        const float axz = d1(&sAz[w], WZ, gx, nx) * a.rdx;  // adjoint z_x
        const float azz = d1(&sAz[w], 1, gz, nz) * a.rdz;   // adjoint z_z
        const float t0 = d1(&sFx[w], WZ, gx, nx) * a.rdx;   // forward x_x
        const float t1 = d1(&sFz[w], 1, gz, nz) * a.rdz;    // forward z_z
        const float t2 = d1(&sFz[w], WZ, gx, nx) * a.rdx;   // forward z_x
        const float t3 = d1(&sFx[w], 1, gz, nz) * a.rdz;    // forward x_z
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
            G[nxz + g] += a.wlm * ((azx + axz) * (t2 + t3) + 2.f * e * axx * t0 + 2.f * azz * t1 + v3 * (axx * t1 + azz * t0));
        }
        G[2 * nxz + g] += a.wrho * (A[g] * d2x + A[nxz + g] * d2z);
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
    switch (tile_x) {
    case 8: hipLaunchKernelGGL((k_step<VTI, 8>), dim3(nb), dim3(kThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_step<VTI, 32>), dim3(nb), dim3(kThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_step<VTI, 16>), dim3(nb), dim3(kThreads), 0, s, a); break;
    }
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
    switch (tile_x) {
    case 8: hipLaunchKernelGGL((k_image<VTI, 8>), dim3(nb), dim3(kThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_image<VTI, 32>), dim3(nb), dim3(kThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_image<VTI, 16>), dim3(nb), dim3(kThreads), 0, s, a); break;
    }
}

void launch_image(const ImageArgs& a, bool vti, int tile_x, hipStream_t s)
{
    if (vti) image_dispatch<true>(a, tile_x, s);
    else image_dispatch<false>(a, tile_x, s);
}

void launch_init_prev(float* prev, long long shot_stride, int nz, int nshots, const int* shot_ptr,
                      const PointEntry* ents, const float* trace, long long comp_stride, hipStream_t s)
{
    // one single-thread block per shot keeps the accumulation order fixed
    hipLaunchKernelGGL(k_init_prev, dim3(nshots), dim3(1), 0, s, prev, shot_stride, nz, shot_ptr, ents, trace,
                       comp_stride, shot_stride / 2);
}

void launch_scale_prev(float* prev, const float* rho, float half_dt2, long long nxz, int nshots, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_prev, dim3(grid_for(2 * nxz * nshots)), dim3(kThreads), 0, s, prev, rho, half_dt2, nxz,
                       nshots);
}

void launch_extract(const float* state, long long shot_stride, int nz, long long nxz, const PointEntry* ents, int n,
                    float* trace, long long comp_stride, int it, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_extract, dim3((2 * n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, state, shot_stride, nz,
                       nxz, ents, n, trace, comp_stride, it);
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
