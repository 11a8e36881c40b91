/*
 * fwi2d oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the reference's elastic (isotropic and VTI)
 * propagation path, used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker.  It is never linked into, or
 * called by, the product library (fwi2d_amd/).
 *
 * Compiled twice from this one source: REAL=double (the FP64 oracle) and
 * REAL=float (used to measure the reference algorithm's own FP32-vs-FP64 gap).
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).  Array layout follows the reference:
 * [..][x][z] with z fastest, index ix*nz+iz.
 *
 * How each part is pinned (details in the tests):
 *   D1 closure + interior        : SBP identity H.D + D^T.H = diag(-1,0..0,1) (exact
 *                                  rationals), polynomial exactness, and the reference's
 *                                  own compiled esat_neumann_top (tangential D1 of row 0)
 *   D2var closure + interior     : reference templates Dzz_var/Dxx_var compiled from source,
 *                                  symmetry of H.D2(c) up to the boundary term, c==1 reduces
 *                                  to the reference's constant-coefficient Dzz closure
 *   SAT neumann/absorbing/Dz/Dx  : reference templates compiled from source
 *   boundary scaling, H          : reference spatial_operators.cpp compiled from source
 *   taper                        : closed form cos^2 values at band ends
 *   delta_m3 injector            : moment conditions (exact for degree <= 2 with H weights)
 *   leapfrog+SBP+free surface    : discrete energy conservation (no taper, no absorbing)
 *   compute_gradients            : finite-difference directional derivative of the L2 misfit
 *   sinc resampler, Dt, Ht       : closed forms / adjoint (dot-product) identities
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* Parameter blocks (fixed layout for both builds; scalars passed as double)  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, nz;
    double dx, dz, ox, oz;
    int nt;            /* propagation steps */
    double dt;         /* propagation time step */
    int sub;           /* snapshot / imaging sub-sampling (0 = none) */
    int bc_top, bc_bottom, bc_left, bc_right;          /* 1 free surface, 2 absorbing */
    int taper_top, taper_bottom, taper_left, taper_right;
    double taper_strength;
    int vti;           /* 0: model = (lambda, mu, rho); 1: (lambda, mu, rho, c13, eps) */
    double gdamp_tmin, gdamp_tmax, gdamp_power;
} orc_grid;

int orc_real_size(void) { return (int)sizeof(real); }

#ifdef _OPENMP
#include <omp.h>
int orc_num_threads(void) { return omp_get_max_threads(); }
#else
int orc_num_threads(void) { return 1; }
#endif

/* Subnormal numbers flushed to zero (x86 MXCSR FTZ and DAZ) on the calling thread and
   every thread of the OpenMP pool.  In single precision the decaying tails of long
   runs become subnormal and slow the CPU many times over; flushing changes values only
   below 1.2e-38.  Returns the number of threads set. */
#if defined(__x86_64__) || defined(__i386__)
#include <xmmintrin.h>
int orc_flush_subnormals(void)
{
    int n = 1;
    _mm_setcsr(_mm_getcsr() | 0x8040);
#ifdef _OPENMP
#pragma omp parallel
    {
        _mm_setcsr(_mm_getcsr() | 0x8040);
#pragma omp single
        n = omp_get_num_threads();
    }
#endif
    return n;
}
#else
int orc_flush_subnormals(void) { return 0; }
#endif

/* ------------------------------------------------------------------------- */
/* SBP coefficient tables                                                      */
/* ------------------------------------------------------------------------- */
/* 4th-order first-derivative closure, rows 0..3 x points 0..5
   (spatial_operators.cpp:102, also the bnd_coef of spatial_operators.hpp:163) */
static const double D1B[4][6] = {
    {-24.0/17, 59.0/34, -4.0/17, -3.0/34, 0, 0},
    {-0.5, 0, 0.5, 0, 0, 0},
    {4.0/43, -59.0/86, 0, 59.0/86, -4.0/43, 0},
    {3.0/98, 0, -59.0/98, 0, 32.0/49, -4.0/49}};
static const double D1C0 = 2.0/3, D1C1 = -1.0/12;         /* spatial_operators.ispc:9-10 */
static const double SCOEF[4] = {11.0/6, -3.0, 1.5, -1.0/3}; /* spatial_operators.hpp:164 */
static const double H0 = 17.0/48;                          /* spatial_operators.hpp:165 */
static const double HCOEF[4] = {17.0/48, 59.0/48, 43.0/48, 49.0/48}; /* spatial_operators.cpp:11 */

/* variable-coefficient second derivative, interior weights (spatial_operators.hpp:34) */
static const double D2C[10] = {-3.0/4, -5.0/6, -1.0/24, 1.0/6, 1.0/2, 1.0/2, 1.0/6, -1.0/8, 1.0/6, -1.0/8};

/* closure: B[r][j][k] for rows r<6, field point j<8, coefficient point k<8
   (spatial_operators.hpp:35-41), stored as its 132 non-zero rationals */
typedef struct { int r, j, k; double num, den; } d2nz;
static const d2nz D2NZ[] = {
    {0,0,0,920.0,289.0}, {0,0,1,-59.0,68.0}, {0,0,2,-7549318.0,34157643.0}, {0,0,3,-17440994.0,184112825.0},
    {0,1,0,-1740.0,289.0}, {0,1,2,295314153.0,366719282.0}, {0,1,3,262545878.0,1218454145.0}, {0,2,0,1128.0,289.0},
    {0,2,1,59.0,68.0}, {0,2,2,-19250923.0,26254840.0}, {0,2,3,-12192537.0,324892213.0}, {0,3,0,-308.0,289.0},
    {0,3,2,42283069.0,254173229.0}, {0,3,3,-43013531.0,427521546.0}, {0,4,2,-18700293.0,1355757959.0}, {0,4,3,18700293.0,1355757959.0},
    {0,5,2,-3.0,833.0}, {0,5,3,3.0,833.0}, {1,0,0,12.0,17.0}, {1,0,2,89562243.0,385991318.0},
    {1,0,3,16914691.0,272440171.0}, {1,1,0,-59.0,68.0}, {1,1,2,-47979680.0,48852831.0}, {1,1,3,-18034913.0,120051851.0},
    {1,2,0,2.0,17.0}, {1,2,2,299262497.0,373256703.0}, {1,2,3,16156647.0,200473586.0}, {1,3,0,3.0,68.0},
    {1,3,2,-14723969.0,177744748.0}, {1,3,3,22633571.0,584543661.0}, {1,4,2,46802031.0,1628311862.0}, {1,4,3,-46802031.0,1628311862.0},
    {1,5,2,441.0,181507.0}, {1,5,3,-441.0,181507.0}, {2,0,0,-96.0,731.0}, {2,0,1,59.0,172.0},
    {2,0,2,-47632751.0,164317197.0}, {2,0,3,-5570723.0,375470930.0}, {2,1,0,118.0,731.0}, {2,1,2,54282598.0,49343777.0},
    {2,1,3,23802793.0,215253532.0}, {2,2,0,-16.0,731.0}, {2,2,1,-59.0,172.0}, {2,2,2,-39119273.0,25083370.0},
    {2,2,3,-35971870.0,61324629.0}, {2,2,4,-26254.0,557679.0}, {2,3,0,-6.0,731.0}, {2,3,2,360454121.0,368940022.0},
    {2,3,3,17254963.0,80047776.0}, {2,3,4,1500708.0,7993399.0}, {2,4,2,-18024731.0,79673021.0}, {2,4,3,24178273.0,88099647.0},
    {2,4,4,-26254.0,185893.0}, {2,5,2,-870707.0,620833782.0}, {2,5,3,960119.0,1147305747.0}, {2,5,4,13564.0,23980197.0},
    {3,0,0,-36.0,833.0}, {3,0,2,54741803.0,948483020.0}, {3,0,3,-13602043.0,389676498.0}, {3,1,0,177.0,3332.0},
    {3,1,2,-35820026.0,359121865.0}, {3,1,3,24921773.0,534548210.0}, {3,2,0,-6.0,833.0}, {3,2,2,813284372.0,948584067.0},
    {3,2,3,30057666.0,158897885.0}, {3,2,4,1500708.0,9108757.0}, {3,3,0,-9.0,3332.0}, {3,3,2,-95056924.0,90903639.0},
    {3,3,3,-23417695.0,47008088.0}, {3,3,4,-7476412.0,9108757.0}, {3,3,5,-2.0,49.0}, {3,4,2,23159719.0,99948527.0},
    {3,4,3,110687545.0,265515812.0}, {3,4,4,4502124.0,9108757.0}, {3,4,5,8.0,49.0}, {3,5,2,-3671038.0,2687426923.0},
    {3,5,3,-1063649.0,8893843.0}, {3,5,4,1473580.0,9108757.0}, {3,5,5,-6.0,49.0}, {4,0,2,-9437957.0,1931986386.0},
    {4,0,3,9437957.0,1931986386.0}, {4,1,2,17289851.0,489388053.0}, {4,1,3,-17289851.0,489388053.0}, {4,2,2,-66355919.0,327412264.0},
    {4,2,3,24178273.0,98343792.0}, {4,2,4,-564461.0,4461432.0}, {4,3,2,17638343.0,74566894.0}, {4,3,3,103749401.0,243793650.0},
    {4,3,4,375177.0,743572.0}, {4,3,5,1.0,6.0}, {4,4,2,-19321801.0,295845927.0}, {4,4,3,-50677283.0,62943042.0},
    {4,4,4,-280535.0,371786.0}, {4,4,5,-5.0,6.0}, {4,4,6,-1.0,24.0}, {4,5,2,5130905.0,5183662662.0},
    {4,5,3,35039615.0,213452232.0}, {4,5,4,1118749.0,2230716.0}, {4,5,5,1.0,2.0}, {4,5,6,1.0,6.0},
    {4,6,4,-1.0,8.0}, {4,6,5,1.0,6.0}, {4,6,6,-1.0,8.0}, {5,0,2,-1.0,784.0},
    {5,0,3,1.0,784.0}, {5,1,2,8673.0,2904112.0}, {5,1,3,-8673.0,2904112.0}, {5,2,2,-403062.0,320810033.0},
    {5,2,3,960119.0,1280713392.0}, {5,2,4,3391.0,6692148.0}, {5,3,2,-1920494.0,1377228165.0}, {5,3,3,-1063649.0,8712336.0},
    {5,3,4,368395.0,2230716.0}, {5,3,5,-1.0,8.0}, {5,4,2,5130905.0,5183662662.0}, {5,4,3,35039615.0,213452232.0},
    {5,4,4,1118749.0,2230716.0}, {5,4,5,1.0,2.0}, {5,4,6,1.0,6.0}, {5,5,2,-117380.0,2351569839.0},
    {5,5,3,-3290636.0,80044587.0}, {5,5,4,-5580181.0,6692148.0}, {5,5,5,-3.0,4.0}, {5,5,6,-5.0,6.0},
    {5,5,7,-1.0,24.0}, {5,6,4,1.0,6.0}, {5,6,5,1.0,2.0}, {5,6,6,1.0,2.0},
    {5,6,7,1.0,6.0}, {5,7,5,-1.0,8.0}, {5,7,6,1.0,6.0}, {5,7,7,-1.0,8.0},
};
static real D2B[6][8][8];
static int d2b_ready = 0;
static void d2b_init(void)
{
    if (d2b_ready) return;
    memset(D2B, 0, sizeof(D2B));
    for (size_t i = 0; i < sizeof(D2NZ) / sizeof(D2NZ[0]); i++)
        D2B[D2NZ[i].r][D2NZ[i].j][D2NZ[i].k] = (real)(D2NZ[i].num / D2NZ[i].den);
    d2b_ready = 1;
}

/* Dense copy of the D2 closure table for tests (r*64 + j*8 + k). */
void orc_d2_table(double* out384)
{
    d2b_init();
    for (int r = 0; r < 6; r++)
        for (int j = 0; j < 8; j++)
            for (int k = 0; k < 8; k++) out384[r * 64 + j * 8 + k] = (double)D2B[r][j][k];
}

/* ------------------------------------------------------------------------- */
/* 1-D line helpers: u[i*s], i in [0,n)                                        */
/* ------------------------------------------------------------------------- */
/* (D1 u)_i * h, optionally of (p*u): closure rows per spatial_operators.cpp:111-119
   (top) / 116-118 (bottom, negated mirror), interior per spatial_operators.ispc:13 */
static inline real d1_line(const real* u, const real* p, long s, int n, int i)
{
    real v = 0;
    if (i < 4) {
        for (int j = 0; j < 6; j++) v += (real)D1B[i][j] * u[j * s] * (p ? p[j * s] : (real)1);
    } else if (i >= n - 4) {
        int r = n - 1 - i;
        for (int j = 0; j < 6; j++) v -= (real)D1B[r][j] * u[(n - 1 - j) * s] * (p ? p[(n - 1 - j) * s] : (real)1);
    } else {
        if (p)
            v = (real)D1C0 * (u[(i + 1) * s] * p[(i + 1) * s] - u[(i - 1) * s] * p[(i - 1) * s]) +
                (real)D1C1 * (u[(i + 2) * s] * p[(i + 2) * s] - u[(i - 2) * s] * p[(i - 2) * s]);
        else
            v = (real)D1C0 * (u[(i + 1) * s] - u[(i - 1) * s]) + (real)D1C1 * (u[(i + 2) * s] - u[(i - 2) * s]);
    }
    return v;
}

/* (D2(c) u)_i * h^2: closure per spatial_operators.hpp:52-76, interior per 79-87 */
static inline real d2_line(const real* u, const real* c, long s, int n, int i)
{
    real out = 0;
    if (i < 6) {
        for (int j = 0; j < 8; j++) {
            real val = 0;
            for (int k = 0; k < 8; k++) val += D2B[i][j][k] * c[k * s];
            out += val * u[j * s];
        }
    } else if (i >= n - 6) {
        int r = n - 1 - i;
        for (int j = 0; j < 8; j++) {
            real val = 0;
            for (int k = 0; k < 8; k++) val += D2B[r][j][k] * c[(n - 1 - k) * s];
            out += val * u[(n - 1 - j) * s];
        }
    } else {
        const real fm2 = c[(i - 2) * s], fm1 = c[(i - 1) * s], f0 = c[i * s], fp1 = c[(i + 1) * s], fp2 = c[(i + 2) * s];
        out = ((real)D2C[0] * f0 + (real)D2C[1] * (fm1 + fp1) + (real)D2C[2] * (fm2 + fp2)) * u[i * s]
            + ((real)D2C[3] * fm2 + (real)D2C[4] * fm1 + (real)D2C[5] * f0 + (real)D2C[6] * fp1) * u[(i - 1) * s]
            + ((real)D2C[3] * fp2 + (real)D2C[4] * fp1 + (real)D2C[5] * f0 + (real)D2C[6] * fm1) * u[(i + 1) * s]
            + ((real)D2C[7] * fm2 + (real)D2C[8] * fm1 + (real)D2C[9] * f0) * u[(i - 2) * s]
            + ((real)D2C[7] * fp2 + (real)D2C[8] * fp1 + (real)D2C[9] * f0) * u[(i + 2) * s];
    }
    return out;
}

/* ------------------------------------------------------------------------- */
/* Whole-grid operators                                                        */
/* ------------------------------------------------------------------------- */
/* out = add*out + a * D_x (p*in) / dx  (Dx: spatial_operators.cpp:128-154; mult_Dx: 184-210) */
void orc_Dx(int add, const real* in, real* out, const real* p, int nx, int nz, double d, double a)
{
    const real sc = (real)(a / d);
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            long o = (long)ix * nz + iz;
            real v = d1_line(in + iz, p ? p + iz : NULL, nz, nx, ix);
            out[o] = (add ? out[o] : (real)0) + sc * v;
        }
}

/* out = add*out + a * D_z (p*in) / dz  (Dz: spatial_operators.cpp:100-126; mult_Dz: 156-182) */
void orc_Dz(int add, const real* in, real* out, const real* p, int nx, int nz, double d, double a)
{
    const real sc = (real)(a / d);
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++) {
        const real* u = in + (long)ix * nz;
        const real* pp = p ? p + (long)ix * nz : NULL;
        for (int iz = 0; iz < nz; iz++) {
            long o = (long)ix * nz + iz;
            real v = d1_line(u, pp, 1, nz, iz);
            out[o] = (add ? out[o] : (real)0) + sc * v;
        }
    }
}

/* out = add*out + a * D2_x(c) in / dx^2  (Dxx_var, spatial_operators.hpp:103-156) */
void orc_Dxx_var(int add, const real* in, real* out, const real* c, int nx, int nz, double d, double a)
{
    d2b_init();
    const real sc = (real)(a / (d * d));
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            long o = (long)ix * nz + iz;
            real v = d2_line(in + iz, c + iz, nz, nx, ix);
            out[o] = (add ? out[o] : (real)0) + sc * v;
        }
}

/* out = add*out + a * D2_z(c) in / dz^2  (Dzz_var, spatial_operators.hpp:31-100) */
void orc_Dzz_var(int add, const real* in, real* out, const real* c, int nx, int nz, double d, double a)
{
    d2b_init();
    const real sc = (real)(a / (d * d));
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            long o = (long)ix * nz + iz;
            real v = d2_line(in + (long)ix * nz, c + (long)ix * nz, 1, nz, iz);
            out[o] = (add ? out[o] : (real)0) + sc * v;
        }
}

/* H quadrature along z / x, in place allowed (applyHz/applyHx, spatial_operators.cpp:6-98) */
void orc_applyH(int axis_x, int inv, real* u, int nx, int nz, double d)
{
    int n = axis_x ? nx : nz;
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            int i = axis_x ? ix : iz;
            double h = (i < 4) ? HCOEF[i] : (i >= n - 4 ? HCOEF[n - 1 - i] : 1.0);
            long o = (long)ix * nz + iz;
            u[o] = inv ? u[o] / (real)(h * d) : u[o] * (real)(h * d);
        }
}

/* ------------------------------------------------------------------------- */
/* SAT boundary terms                                                          */
/* ------------------------------------------------------------------------- */
/* Free-surface / absorbing SAT on one boundary line (spatial_operators.hpp:158-727).
   side: 0 top, 1 bottom, 2 left, 3 right.  in0 is differentiated along the boundary,
   in1 along the outward normal (one-sided SCOEF), in2 (may be NULL) is the previous
   time level for the absorbing term with coefficient f3.  out += SAT. */
void orc_esat(int side, const real* in0, const real* in1, const real* in2, real* out,
              const real* f1, const real* f2, const real* f3, int nx, int nz, double dx, double dz, double dt)
{
    if (side <= 1) {
        const double hn = dz, ht = dx;
        const double sgn = side == 0 ? -1.0 : 1.0; /* top: -f1, bottom: +f1 */
        for (int ix = 0; ix < nx; ix++) {
            int iz0 = side == 0 ? 0 : nz - 1;
            long o = (long)ix * nz + iz0;
            real sumx = d1_line(in0 + iz0, NULL, nz, nx, ix);
            real sumz = 0;
            for (int k = 0; k < 4; k++) sumz += (real)SCOEF[k] * in1[(long)ix * nz + (side == 0 ? k : nz - 1 - k)];
            real term = (real)sgn * f1[o] * sumx / (real)ht + f2[o] * sumz / (real)hn;
            if (in2) term -= f3[o] * in2[o] / (real)dt;
            out[o] -= (real)(1.0 / (hn * H0)) * term;
        }
    } else {
        const double hn = dx, ht = dz;
        const double sgn = side == 2 ? -1.0 : 1.0; /* left: -f1, right: +f1 */
        int ix0 = side == 2 ? 0 : nx - 1;
        for (int iz = 0; iz < nz; iz++) {
            long o = (long)ix0 * nz + iz;
            real sumz = d1_line(in0 + (long)ix0 * nz, NULL, 1, nz, iz);
            real sumx = 0;
            for (int k = 0; k < 4; k++) sumx += (real)SCOEF[k] * in1[(long)(side == 2 ? k : nx - 1 - k) * nz + iz];
            real term = (real)sgn * f1[o] * sumz / (real)ht + f2[o] * sumx / (real)hn;
            if (in2) term -= f3[o] * in2[o] / (real)dt;
            out[o] -= (real)(1.0 / (hn * H0)) * term;
        }
    }
}

/* Complement of the wide-stencil operator on the boundary line:
   out_b += (+/-) f * (D_n in)_b / (h*h0)  (esat_Dz_top/bottom, esat_Dx_left/right,
   spatial_operators.hpp:730-822) */
void orc_esat_dn(int side, const real* in, real* out, const real* f, int nx, int nz, double dx, double dz)
{
    if (side <= 1) {
        const real sc = (real)(1.0 / (dz * dz * H0));
        for (int ix = 0; ix < nx; ix++) {
            real sum = 0;
            if (side == 0) { for (int j = 0; j < 6; j++) sum -= (real)D1B[0][j] * in[(long)ix * nz + j]; }
            else           { for (int j = 0; j < 6; j++) sum -= (real)D1B[0][j] * in[(long)ix * nz + nz - 1 - j]; }
            long o = (long)ix * nz + (side == 0 ? 0 : nz - 1);
            out[o] -= sc * f[o] * sum;
        }
    } else {
        const real sc = (real)(1.0 / (dx * dx * H0));
        for (int iz = 0; iz < nz; iz++) {
            real sum = 0;
            if (side == 2) { for (int j = 0; j < 6; j++) sum -= (real)D1B[0][j] * in[(long)j * nz + iz]; }
            else           { for (int j = 0; j < 6; j++) sum -= (real)D1B[0][j] * in[(long)(nx - 1 - j) * nz + iz]; }
            long o = (long)(side == 2 ? 0 : nx - 1) * nz + iz;
            out[o] -= sc * f[o] * sum;
        }
    }
}

/* Implicit part of the absorbing SAT: divide boundary values of the new level by
   1 + c.dt/(2 h h0), accumulating corner contributions exactly as
   esat_scale_boundaries / vtisat_scale_boundaries (spatial_operators.cpp:307-521)
   do for the full-grid call (ixmin=0, ixmax=nx, izmin=0, izmax=nz).  For VTI the
   x-directed P velocity uses (1+2eps)(lambda+2mu).  model = (lambda, mu, rho[, c13, eps]). */
void orc_scale_boundaries(real* ux, real* uz, const real* model, int vti, int nx, int nz,
                          double dx, double dz, double dt, int top, int bottom, int left, int right)
{
    const long nxz = (long)nx * nz;
    const real* la = model;
    const real* mu = model + nxz;
    const real* rho = model + 2 * nxz;
    const real* eps = vti ? model + 4 * nxz : NULL;
    const real h0 = (real)H0, rdt = (real)dt, rdx = (real)dx, rdz = (real)dz;
#define VPX2(i) ((vti ? (1 + 2 * eps[i]) : (real)1) * (la[i] + 2 * mu[i]))
    real sc1x = 0, sc1z = 0, sc2x = 0, sc2z = 0, sx, sz;
    if (top) {
        sc1x = (real)sqrt(mu[0] / rho[0]) * rdt / (2 * rdz * h0);
        sc1z = sc1x * (real)sqrt((la[0] + 2 * mu[0]) / mu[0]);
        long tr = (long)(nx - 1) * nz;
        sc2x = (real)sqrt(mu[tr] / rho[tr]) * rdt / (2 * rdz * h0);
        sc2z = sc1x * (real)sqrt((la[tr] + 2 * mu[tr]) / mu[tr]); /* uses sc1x as the reference does */
        for (int ix = 1; ix < nx - 1; ix++) {
            long i = (long)ix * nz;
            sx = (real)sqrt(mu[i] / rho[i]) * rdt / (2 * rdz * h0);
            sz = sx * (real)sqrt((la[i] + 2 * mu[i]) / mu[i]);
            ux[i] = ux[i] / (1 + sx);
            uz[i] = uz[i] / (1 + sz);
        }
    }
    if (left) {
        sc1x += (real)sqrt(VPX2(0) / rho[0]) * rdt / (2 * rdx * h0);
        sc1z += (real)sqrt(mu[0] / rho[0]) * rdt / (2 * rdx * h0);
        ux[0] = ux[0] / (1 + sc1x);
        uz[0] = uz[0] / (1 + sc1z);
        long bl = nz - 1;
        sc1x = (real)sqrt(VPX2(bl) / rho[bl]) * rdt / (2 * rdx * h0);
        sc1z = sc1x * (real)sqrt(mu[bl] / VPX2(bl));
        for (int iz = 1; iz < nz - 1; iz++) {
            sx = (real)sqrt(VPX2(iz) / rho[iz]) * rdt / (2 * rdx * h0);
            sz = sx * (real)sqrt(mu[iz] / VPX2(iz));
            ux[iz] = ux[iz] / (1 + sx);
            uz[iz] = uz[iz] / (1 + sz);
        }
    } else {
        ux[0] = ux[0] / (1 + sc1x);
        uz[0] = uz[0] / (1 + sc1z);
        sc1x = 0;
        sc1z = 0;
    }
    if (bottom) {
        long bl = nz - 1;
        sc1x += (real)sqrt(mu[bl] / rho[bl]) * rdt / (2 * rdz * h0);
        sc1z += (real)sqrt((la[bl] + 2 * mu[bl]) / rho[bl]) * rdt / (2 * rdz * h0);
        ux[bl] = ux[bl] / (1 + sc1x);
        uz[bl] = uz[bl] / (1 + sc1z);
        long br = (long)(nx - 1) * nz + nz - 1;
        sc1x = (real)sqrt(mu[br] / rho[br]) * rdt / (2 * rdz * h0);
        sc1z = sc1x * (real)sqrt((la[br] + 2 * mu[br]) / mu[br]);
        for (int ix = 1; ix < nx - 1; ix++) {
            long i = (long)ix * nz + nz - 1;
            sx = (real)sqrt(mu[i] / rho[i]) * rdt / (2 * rdz * h0);
            sz = sx * (real)sqrt((la[i] + 2 * mu[i]) / mu[i]);
            ux[i] = ux[i] / (1 + sx);
            uz[i] = uz[i] / (1 + sz);
        }
    } else {
        ux[nz - 1] = ux[nz - 1] / (1 + sc1x);
        uz[nz - 1] = uz[nz - 1] / (1 + sc1z);
        sc1x = 0;
        sc1z = 0;
    }
    if (right) {
        long br = (long)(nx - 1) * nz + nz - 1;
        sc1x += (real)sqrt(VPX2(br) / rho[br]) * rdt / (2 * rdx * h0);
        sc1z += (real)sqrt(mu[br] / rho[br]) * rdt / (2 * rdx * h0);
        ux[br] = ux[br] / (1 + sc1x);
        uz[br] = uz[br] / (1 + sc1z);
        for (int iz = 1; iz < nz - 1; iz++) {
            long i = (long)(nx - 1) * nz + iz;
            sx = (real)sqrt(VPX2(i) / rho[i]) * rdt / (2 * rdx * h0);
            sz = sx * (real)sqrt(mu[i] / VPX2(i));
            ux[i] = ux[i] / (1 + sx);
            uz[i] = uz[i] / (1 + sz);
        }
        long tr = (long)(nx - 1) * nz;
        sc2x += (real)sqrt(VPX2(tr) / rho[tr]) * rdt / (2 * rdx * h0);
        sc2z += (real)sqrt(mu[tr] / rho[tr]) * rdt / (2 * rdx * h0);
        ux[tr] = ux[tr] / (1 + sc2x);
        uz[tr] = uz[tr] / (1 + sc2z);
    } else {
        long br = (long)(nx - 1) * nz + nz - 1;
        ux[br] = ux[br] / (1 + sc1x);
        uz[br] = uz[br] / (1 + sc1z);
        long tr = (long)(nx - 1) * nz;
        ux[tr] = ux[tr] / (1 + sc2x);
        uz[tr] = uz[tr] / (1 + sc2z);
    }
#undef VPX2
}

/* cos^2 taper on the four bands (taperz/taperx, spatial_operators.cpp:523-549,
   spatial_operators.ispc:47-69, called as at we_op.cpp:846-861) */
static inline real taper_factor(int i, int n, int t_lo, int t_hi, double a)
{
    real f = 1;
    if (t_lo > 0 && i < t_lo) {
        real v = (real)cos(a * 0.5 * M_PI * (double)(i + 1 - t_lo) / (double)(0 - t_lo));
        f *= v * v;
    }
    if (t_hi > 0 && i >= n - t_hi) {
        real v = (real)cos(a * 0.5 * M_PI * (double)(i - (n - t_hi)) / (double)t_hi);
        f *= v * v;
    }
    return f;
}

void orc_taper(real* u, int nx, int nz, int ttop, int tbot, int tleft, int tright, double a)
{
#pragma omp parallel for
    for (int ix = 0; ix < nx; ix++) {
        real fx = taper_factor(ix, nx, tleft, tright, a);
        for (int iz = 0; iz < nz; iz++) {
            real fz = taper_factor(iz, nz, ttop, tbot, a);
            long o = (long)ix * nz + iz;
            /* reference order: z bands first (top, bottom), then x bands (left, right) */
            u[o] *= fz;
            u[o] *= fx;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Point injector / extractor: delta_m3 (injector.cpp:3-112)                   */
/* ------------------------------------------------------------------------- */
static double hcoef_at(int n, int i)   /* getHcoef, misc.cpp:11-16 */
{
    if (i < 4) return HCOEF[i];
    if (i >= n - 4) return HCOEF[n - 1 - i];
    return 1.0;
}

/* xind/zind: [ntr][3]; xw/zw: [ntr][2][3] (row 0 injection, row 1 extraction) */
void orc_delta_m3(int nx, int nz, double dx, double dz, double ox, double oz, int ntr,
                  const double* locx, const double* locz, int* xind, int* zind, real* xw, real* zw)
{
    for (int t = 0; t < ntr; t++) {
        int px = (int)lround((locx[t] - ox) / dx);
        int pz = (int)lround((locz[t] - oz) / dz);
        if (px == 0) px = 1;
        if (px == nx - 1) px = nx - 2;
        if (pz == 0) pz = 1;
        if (pz == nz - 1) pz = nz - 2;
        for (int k = 0; k < 3; k++) { xind[3 * t + k] = px - 1 + k; zind[3 * t + k] = pz - 1 + k; }
        real ax = (real)((locx[t] - dx * px - ox) / dx);
        real az = (real)((locz[t] - dz * pz - oz) / dz);
        real* wx = xw + 6 * t;
        real* wz = zw + 6 * t;
        wx[0] = (ax * ax - ax) / (real)(2 * dx * hcoef_at(nx, px - 1));
        wx[1] = (-ax * ax + 1) / (real)(dx * hcoef_at(nx, px));
        wx[2] = (ax * ax + ax) / (real)(2 * dx * hcoef_at(nx, px + 1));
        wz[0] = (az * az - az) / (real)(2 * dz * hcoef_at(nz, pz - 1));
        wz[1] = (-az * az + 1) / (real)(dz * hcoef_at(nz, pz));
        wz[2] = (az * az + az) / (real)(2 * dz * hcoef_at(nz, pz + 1));
        wx[3] = (ax * ax - ax) / 2; wx[4] = -ax * ax + 1; wx[5] = (ax * ax + ax) / 2;
        wz[3] = (az * az - az) / 2; wz[4] = -az * az + 1; wz[5] = (az * az + az) / 2;
    }
}

/* ddelta_m3 (injector.cpp:114-239): the delta_m3 footprint of each point with, in
   columns 3..5, the derivative weights: injection (1-2a)/(2h^2 H), 2a/(h^2 H),
   (-1-2a)/(2h^2 H); extraction (1-2a)/(2h), 2a/h, (-1-2a)/(2h).
   xind/zind: [ntr][3]; xw/zw: [ntr][2][6] (row 0 injection, row 1 extraction). */
void orc_ddelta_m3(int nx, int nz, double dx, double dz, double ox, double oz, int ntr,
                   const double* locx, const double* locz, int* xind, int* zind, real* xw, real* zw)
{
    for (int t = 0; t < ntr; t++) {
        int px = (int)lround((locx[t] - ox) / dx);
        int pz = (int)lround((locz[t] - oz) / dz);
        if (px == 0) px = 1;
        if (px == nx - 1) px = nx - 2;
        if (pz == 0) pz = 1;
        if (pz == nz - 1) pz = nz - 2;
        for (int k = 0; k < 3; k++) { xind[3 * t + k] = px - 1 + k; zind[3 * t + k] = pz - 1 + k; }
        real ax = (real)((locx[t] - dx * px - ox) / dx);
        real az = (real)((locz[t] - dz * pz - oz) / dz);
        real* wx = xw + 12 * t;
        real* wz = zw + 12 * t;
        const double hx[3] = {hcoef_at(nx, px - 1), hcoef_at(nx, px), hcoef_at(nx, px + 1)};
        const double hz[3] = {hcoef_at(nz, pz - 1), hcoef_at(nz, pz), hcoef_at(nz, pz + 1)};
        wx[0] = (ax * ax - ax) / (real)(2 * dx * hx[0]);
        wx[1] = (-ax * ax + 1) / (real)(dx * hx[1]);
        wx[2] = (ax * ax + ax) / (real)(2 * dx * hx[2]);
        wx[3] = (1 - 2 * ax) / (real)(2 * dx * dx * hx[0]);
        wx[4] = 2 * ax / (real)(dx * dx * hx[1]);
        wx[5] = (-1 - 2 * ax) / (real)(2 * dx * dx * hx[2]);
        wz[0] = (az * az - az) / (real)(2 * dz * hz[0]);
        wz[1] = (-az * az + 1) / (real)(dz * hz[1]);
        wz[2] = (az * az + az) / (real)(2 * dz * hz[2]);
        wz[3] = (1 - 2 * az) / (real)(2 * dz * dz * hz[0]);
        wz[4] = 2 * az / (real)(dz * dz * hz[1]);
        wz[5] = (-1 - 2 * az) / (real)(2 * dz * dz * hz[2]);
        wx[6] = (ax * ax - ax) / 2; wx[7] = -ax * ax + 1; wx[8] = (ax * ax + ax) / 2;
        wx[9] = (1 - 2 * ax) / (real)(2 * dx); wx[10] = 2 * ax / (real)dx; wx[11] = (-1 - 2 * ax) / (real)(2 * dx);
        wz[6] = (az * az - az) / 2; wz[7] = -az * az + 1; wz[8] = (az * az + az) / 2;
        wz[9] = (1 - 2 * az) / (real)(2 * dz); wz[10] = 2 * az / (real)dz; wz[11] = (-1 - 2 * az) / (real)(2 * dz);
    }
}

/* dipole_m3 (injector.cpp:241-378): two delta_m3 footprints at the ends of a gauge of length
   gl centred on each point along its dip (x -/+ gl/2 cos(dip), z -/+ gl/2 sin(dip)).
   xind/zind: [ntr][6] (columns 0..2 first end, 3..5 second); xw/zw: [ntr][2][6]. */
void orc_dipole_m3(int nx, int nz, double dx, double dz, double ox, double oz, int ntr, const double* locx,
                   const double* locz, const double* dip, double gl, int* xind, int* zind, real* xw, real* zw)
{
    for (int t = 0; t < ntr; t++) {
        const real c = (real)cos(dip[t]), sn = (real)sin(dip[t]);
        const real hg = (real)(gl / 2);
        double ex[2] = {locx[t] - hg * c, locx[t] + hg * c};
        double ez[2] = {locz[t] - hg * sn, locz[t] + hg * sn};
        for (int e = 0; e < 2; e++) {
            int px = (int)lround((ex[e] - ox) / dx);
            int pz = (int)lround((ez[e] - oz) / dz);
            if (px == 0) px = 1;
            if (px == nx - 1) px = nx - 2;
            if (pz == 0) pz = 1;
            if (pz == nz - 1) pz = nz - 2;
            for (int k = 0; k < 3; k++) { xind[6 * t + 3 * e + k] = px - 1 + k; zind[6 * t + 3 * e + k] = pz - 1 + k; }
            real ax = (real)((ex[e] - ox) / dx - px);
            real az = (real)((ez[e] - oz) / dz - pz);
            real* wx = xw + 12 * t + 3 * e;
            real* wz = zw + 12 * t + 3 * e;
            wx[0] = (ax * ax - ax) / (real)(2 * dx * hcoef_at(nx, px - 1));
            wx[1] = (-ax * ax + 1) / (real)(dx * hcoef_at(nx, px));
            wx[2] = (ax * ax + ax) / (real)(2 * dx * hcoef_at(nx, px + 1));
            wz[0] = (az * az - az) / (real)(2 * dz * hcoef_at(nz, pz - 1));
            wz[1] = (-az * az + 1) / (real)(dz * hcoef_at(nz, pz));
            wz[2] = (az * az + az) / (real)(2 * dz * hcoef_at(nz, pz + 1));
            wx[6] = (ax * ax - ax) / 2; wx[7] = -ax * ax + 1; wx[8] = (ax * ax + ax) / 2;
            wz[6] = (az * az - az) / 2; wz[7] = -az * az + 1; wz[8] = (az * az + az) / 2;
        }
    }
}

/* dipole_to_strain (misc.cpp:18-43) on [2][ntr][nt]: forward, p0 = (cos p0 + sin p1)/gl
   and p1 = 0; adjoint, p1 = sin p0/gl and p0 = cos p0/gl. */
void orc_dipole_to_strain(int adj, real* p, const double* dip, int ntr, int nt, double gl)
{
    for (int x = 0; x < ntr; x++) {
        const real c = (real)cos(dip[x]), sn = (real)sin(dip[x]);
        real* p0 = p + (long)x * nt;
        real* p1 = p + (long)(ntr + x) * nt;
        for (int t = 0; t < nt; t++) {
            if (!adj) { p0[t] = (c * p0[t] + sn * p1[t]) / (real)gl; p1[t] = 0; }
            else { p1[t] = sn * p0[t] / (real)gl; p0[t] *= c / (real)gl; }
        }
    }
}

/* A point injector or extractor as lists of (grid cell, weight, trace row) per field
   component (0 = x, 1 = z): injection adds w * row[it] to the cell, extraction adds w times
   the cell to row[it]; rows are offsets from a base pointer per component.  delta_m3,
   ddelta_m3 and dipole_m3 are all of this form; the entries follow the reference's loop
   order (trace, then x, then z offset of the footprint). */
typedef struct {
    int n[2];
    long* cell[2];
    long* off[2];
    real* w[2];
} pts;

static void pts_alloc(pts* p, int n)
{
    for (int c = 0; c < 2; c++) {
        p->n[c] = 0;
        p->cell[c] = (long*)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
        p->off[c] = (long*)malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
        p->w[c] = (real*)malloc(sizeof(real) * (size_t)(n > 0 ? n : 1));
    }
}

static void pts_free(pts* p)
{
    for (int c = 0; c < 2; c++) { free(p->cell[c]); free(p->off[c]); free(p->w[c]); }
}

static void pts_add(pts* p, int c, long cell, long off, real w)
{
    const int k = p->n[c]++;
    p->cell[c][k] = cell; p->off[c][k] = off; p->w[c][k] = w;
}

/* delta_m3 footprints, both components reading (writing) trace t at offset t * nt from
   their base; row = 0 injection weights, 1 extraction */
static void pts_delta(pts* p, int ntr, const int* xi, const int* zi, const real* xw, const real* zw, int row,
                      int nt, int nz)
{
    pts_alloc(p, 9 * ntr);
    for (int c = 0; c < 2; c++)
        for (int t = 0; t < ntr; t++)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    pts_add(p, c, (long)xi[3 * t + a] * nz + zi[3 * t + b], (long)t * nt,
                            xw[6 * t + 3 * row + a] * zw[6 * t + 3 * row + b]);
}

/* ddelta_m3 injection of a moment tensor (we_op.cpp:640-654; injector.cpp:191-217): the
   x component takes -d/dx of Mxx and -d/dz of Mxz, the z component -d/dx of Mxz and -d/dz
   of Mzz; offsets of the Mxx, Mzz and Mxz traces of point t are oxx + t nt, ozz + t nt,
   oxz + t nt from one base */
static void pts_ddelta(pts* p, int ntr, const int* xi, const int* zi, const real* xw, const real* zw, long oxx,
                       long ozz, long oxz, int nt, int nz)
{
    pts_alloc(p, 18 * ntr);
    for (int t = 0; t < ntr; t++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const long cell = (long)xi[3 * t + a] * nz + zi[3 * t + b];
                const real wdx = -xw[12 * t + 3 + a] * zw[12 * t + b];
                const real wdz = -xw[12 * t + a] * zw[12 * t + 3 + b];
                pts_add(p, 0, cell, oxx + (long)t * nt, wdx);
                pts_add(p, 0, cell, oxz + (long)t * nt, wdz);
                pts_add(p, 1, cell, oxz + (long)t * nt, wdx);
                pts_add(p, 1, cell, ozz + (long)t * nt, wdz);
            }
}

/* dipole_m3 (injector.cpp:333-378): -first end + second end, both components on trace t */
static void pts_dipole(pts* p, int ntr, const int* xi, const int* zi, const real* xw, const real* zw, int row,
                       int nt, int nz)
{
    pts_alloc(p, 18 * ntr);
    for (int c = 0; c < 2; c++)
        for (int t = 0; t < ntr; t++)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    if (row == 0) {  /* inject: the first end subtracts, the second adds */
                        pts_add(p, c, (long)xi[6 * t + a] * nz + zi[6 * t + b], (long)t * nt,
                                -xw[12 * t + a] * zw[12 * t + b]);
                        pts_add(p, c, (long)xi[6 * t + 3 + a] * nz + zi[6 * t + 3 + b], (long)t * nt,
                                xw[12 * t + 3 + a] * zw[12 * t + 3 + b]);
                    } else {  /* extract: second end minus first end */
                        pts_add(p, c, (long)xi[6 * t + 3 + a] * nz + zi[6 * t + 3 + b], (long)t * nt,
                                xw[12 * t + 9 + a] * zw[12 * t + 9 + b]);
                        pts_add(p, c, (long)xi[6 * t + a] * nz + zi[6 * t + b], (long)t * nt,
                                -xw[12 * t + 6 + a] * zw[12 * t + 6 + b]);
                    }
                }
}

static void pts_inject(const pts* p, const real* base0, const real* base1, int it, real* ux, real* uz)
{
    const real* base[2] = {base0, base1};
    real* u[2] = {ux, uz};
    for (int c = 0; c < 2; c++)
        for (int k = 0; k < p->n[c]; k++) u[c][p->cell[c][k]] += p->w[c][k] * base[c][p->off[c][k] + it];
}

static void pts_extract(const pts* p, real* base0, real* base1, int it, const real* ux, const real* uz)
{
    real* base[2] = {base0, base1};
    const real* u[2] = {ux, uz};
    for (int c = 0; c < 2; c++)
        for (int k = 0; k < p->n[c]; k++) base[c][p->off[c][k] + it] += p->w[c][k] * u[c][p->cell[c][k]];
}

/* ------------------------------------------------------------------------- */
/* Imaging condition                                                           */
/* ------------------------------------------------------------------------- */
/* nl_we_op_e::compute_gradients (we_op.cpp:505-562) and nl_we_op_vti::compute_gradients
   (we_op.cpp:1478-1557).  full = forward snapshots [nsnap][2][nx][nz]; adj = current
   adjoint level [2][nx][nz]; ax = (Dx adj_x, Dx adj_z) and az = (Dz adj_x, Dz adj_z) of the
   previous adjoint level; grad [3|5][nx][nz] accumulated.  dtg = sub*dt. */
static void compute_gradients(const orc_grid* g, const real* model, const real* full, const real* adj,
                              const real* ax, const real* az, real* tmp, real* grad, int idx, double dtg)
{
    const int nx = g->nx, nz = g->nz;
    const long nxz = (long)nx * nz;
    const real* F = full + (long)idx * 2 * nxz;
    const real* Fn = full + (long)(idx + 1) * 2 * nxz;
    const real* Fp = full + (long)(idx - 1) * 2 * nxz;
    real* t0 = tmp;            /* forward x_x */
    real* t1 = tmp + nxz;      /* forward z_z */
    real* t2 = tmp + 2 * nxz;  /* forward z_x */
    real* t3 = tmp + 3 * nxz;  /* forward x_z */
    orc_Dx(0, F, t0, NULL, nx, nz, g->dx, 1.0);
    orc_Dz(0, F + nxz, t1, NULL, nx, nz, g->dz, 1.0);
    orc_Dx(0, F + nxz, t2, NULL, nx, nz, g->dx, 1.0);
    orc_Dz(0, F, t3, NULL, nx, nz, g->dz, 1.0);

    double t = idx * dtg;
    double w = 1;
    if (g->gdamp_tmax - t > 0 && t - g->gdamp_tmin >= 0)
        w = cos(g->gdamp_power * M_PI * (t - g->gdamp_tmin) / (g->gdamp_tmax - g->gdamp_tmin));
    w *= w;
    const int last = g->nt / g->sub;
    const real dt = (real)dtg, rw = (real)w;
    const real wlm = (idx == 0) ? rw * (real)0.5 * dt : rw * dt; /* lambda/mu weight */
    const real* la = model;
    const real* mu = model + nxz;
    const real* c13 = model + 3 * nxz;
    const real* eps = model + 4 * nxz;
    const real* axx = ax;           /* Dx adj_x */
    const real* axz = ax + nxz;     /* Dx adj_z */
    const real* azx = az;           /* Dz adj_x */
    const real* azz = az + nxz;     /* Dz adj_z */
#pragma omp parallel for
    for (long i = 0; i < nxz; i++) {
        real d2x, d2z;
        if (idx > 0 && idx < last) {
            d2x = Fn[i] - 2 * F[i] + Fp[i];
            d2z = Fn[nxz + i] - 2 * F[nxz + i] + Fp[nxz + i];
        } else if (idx == 0) {
            d2x = Fn[i] - F[i];
            d2z = Fn[nxz + i] - F[nxz + i];
        } else {
            d2x = -F[i] + Fp[i];
            d2z = -F[nxz + i] + Fp[nxz + i];
        }
        if (!g->vti) {
            grad[i] += wlm * (axx[i] + azz[i]) * (t0[i] + t1[i]);
            grad[nxz + i] += wlm * ((azx[i] + axz[i]) * (t2[i] + t3[i]) + 2 * axx[i] * t0[i] + 2 * azz[i] * t1[i]);
        } else {
            real l = la[i], m = mu[i];
            real del = ((c13[i] + m) * (c13[i] + m) - (l + m) * (l + m)) / (2 * (l + m) * (l + 2 * m));
            real v1 = (real)sqrt(2 * (l + 2 * m) * (l + m) * del + (l + m) * (l + m));
            real v2 = ((1 + 2 * del) * l + (1 + 3 * del) * m) / v1;
            real v3 = ((1 + 3 * del) * l + (1 + 4 * del) * m) / v1 - 1;
            real e = 1 + 2 * eps[i];
            grad[i] += wlm * (e * axx[i] * t0[i] + azz[i] * t1[i] + v2 * (axx[i] * t1[i] + azz[i] * t0[i]));
            grad[nxz + i] += wlm * ((azx[i] + axz[i]) * (t2[i] + t3[i]) + 2 * e * axx[i] * t0[i] + 2 * azz[i] * t1[i]
                                    + v3 * (axx[i] * t1[i] + azz[i] * t0[i]));
            grad[3 * nxz + i] = 0;
            grad[4 * nxz + i] = 0;
        }
        grad[2 * nxz + i] += rw / dt * (adj[i] * d2x + adj[nxz + i] * d2z);
    }
}

/* ------------------------------------------------------------------------- */
/* Time loop: nl_we_op_e::propagate (we_op.cpp:564-902), nl_we_op_vti::propagate   */
/* (we_op.cpp:1559-1861), version 2 operators, no PML.                          */
/* ------------------------------------------------------------------------- */
/* Coefficient ("expression") fields of we_op.cpp:497-503 (iso) and 1465-1476 (vti). */
typedef struct {
    real *f_2mu_x, *f_2mu_z, *f_imp_s, *f_imp_px, *f_imp_pz, *f_lr_f1, *f_lr_f2, *f_dxc, *f_tb_f1;
} coef_fields;

static void coef_init(const orc_grid* g, const real* model, real* buf, coef_fields* c)
{
    const long nxz = (long)g->nx * g->nz;
    const int vti = g->vti;
    c->f_2mu_x = buf;
    c->f_2mu_z = buf + nxz;
    c->f_imp_s = buf + 2 * nxz;
    c->f_imp_px = buf + 3 * nxz;   /* P impedance for left/right (x-directed) */
    c->f_imp_pz = buf + 4 * nxz;   /* P impedance for top/bottom */
    c->f_lr_f1 = buf + 5 * nxz;    /* lambda (iso) or c13 (vti): tangential coefficient, left/right */
    c->f_lr_f2 = buf + 6 * nxz;    /* 2mu (iso) or (1+2eps)2mu (vti) */
    c->f_dxc = buf + 7 * nxz;      /* lambda (iso) or (1+2eps)lambda (vti): esat_Dx coefficient */
    c->f_tb_f1 = buf + 8 * nxz;    /* lambda (iso) or c13 (vti): tangential coefficient, top/bottom */
    const real* la = model;
    const real* mu = model + nxz;
    const real* rho = model + 2 * nxz;
    const real* c13 = vti ? model + 3 * nxz : NULL;
    const real* eps = vti ? model + 4 * nxz : NULL;
    for (long i = 0; i < nxz; i++) {
        c->f_2mu_z[i] = 2 * mu[i];
        c->f_2mu_x[i] = vti ? (1 + 2 * eps[i]) * 2 * mu[i] : 2 * mu[i];
        c->f_imp_s[i] = (real)(0.5 * sqrt(mu[i] * rho[i]));
        c->f_imp_pz[i] = (real)(0.5 * sqrt((la[i] + 2 * mu[i]) * rho[i]));
        c->f_imp_px[i] = vti ? (real)(0.5 * sqrt((1 + 2 * eps[i]) * (la[i] + 2 * mu[i]) * rho[i])) : c->f_imp_pz[i];
        c->f_lr_f1[i] = vti ? c13[i] : la[i];
        c->f_tb_f1[i] = vti ? c13[i] : la[i];
        c->f_lr_f2[i] = c->f_2mu_x[i];
        c->f_dxc[i] = vti ? (1 + 2 * eps[i]) * la[i] : la[i];
    }
}

/* Volume terms, version 2 (we_op.cpp:690-720 iso, 1665-1700 vti).  Writes the four
   first derivatives of curr into ux = (Dx u_x, Dx u_z), uzd = (Dz u_x, Dz u_z). */
static void rhs_volume(const orc_grid* g, const real* model, const coef_fields* c, const real* cx, const real* cz,
                       real* nxt, real* nzt, real* ux, real* uzd, real* tmp)
{
    const int nx = g->nx, nz = g->nz, vti = g->vti;
    const long nxz = (long)nx * nz;
    const double dx = g->dx, dz = g->dz;
    const real* la = model;
    const real* mu = model + nxz;
    const real* c13 = vti ? model + 3 * nxz : NULL;
    orc_Dxx_var(0, cx, nxt, c->f_2mu_x, nx, nz, dx, 1.0);
    orc_Dzz_var(0, cz, nzt, c->f_2mu_z, nx, nz, dz, 1.0);
    orc_Dxx_var(1, cz, nzt, mu, nx, nz, dx, 1.0);
    orc_Dzz_var(1, cx, nxt, mu, nx, nz, dz, 1.0);
    orc_Dx(0, cx, ux, NULL, nx, nz, dx, 1.0);          /* u_x[0] = Dx ux */
    orc_Dz(0, cz, uzd + nxz, NULL, nx, nz, dz, 1.0);   /* u_z[1] = Dz uz */
    if (!vti) {
        for (long i = 0; i < nxz; i++) tmp[i] = ux[i] + uzd[nxz + i];
        orc_Dx(1, tmp, nxt, la, nx, nz, dx, 1.0);
        orc_Dz(1, tmp, nzt, la, nx, nz, dz, 1.0);
    } else {
        for (long i = 0; i < nxz; i++) tmp[i] = c->f_dxc[i] * ux[i] + c13[i] * uzd[nxz + i];
        orc_Dx(1, tmp, nxt, NULL, nx, nz, dx, 1.0);
        for (long i = 0; i < nxz; i++) tmp[i] = la[i] * uzd[nxz + i] + c13[i] * ux[i];
        orc_Dz(1, tmp, nzt, NULL, nx, nz, dz, 1.0);
    }
    orc_Dx(0, cz, ux + nxz, NULL, nx, nz, dx, 1.0);    /* u_x[1] = Dx uz */
    orc_Dz(0, cx, uzd, NULL, nx, nz, dz, 1.0);         /* u_z[0] = Dz ux */
    orc_Dx(1, uzd, nzt, mu, nx, nz, dx, 1.0);          /* next_z += Dx(mu Dz ux) */
    orc_Dz(1, ux + nxz, nxt, mu, nx, nz, dz, 1.0);     /* next_x += Dz(mu Dx uz) */
}

/* Boundary SAT terms (we_op.cpp:726-822 iso, 1705-1800 vti). px/pz: previous level
   (used by the absorbing terms only). */
static void rhs_sat(const orc_grid* g, const real* model, const coef_fields* c, const real* cx, const real* cz,
                    const real* px, const real* pz, real* nxt, real* nzt)
{
    const int nx = g->nx, nz = g->nz;
    const long nxz = (long)nx * nz;
    const double dx = g->dx, dz = g->dz, dt = g->dt;
    const real* la = model;
    const real* mu = model + nxz;
    for (int side = 0; side < 2; side++) {
        int bc = side == 0 ? g->bc_top : g->bc_bottom;
        if (bc != 1 && bc != 2) continue;
        orc_esat(side, cz, cx, bc == 2 ? px : NULL, nxt, mu, mu, c->f_imp_s, nx, nz, dx, dz, dt);
        orc_esat(side, cx, cz, bc == 2 ? pz : NULL, nzt, c->f_tb_f1, c->f_2mu_z, c->f_imp_pz, nx, nz, dx, dz, dt);
        orc_esat_dn(side, cz, nzt, la, nx, nz, dx, dz);
    }
    for (int side = 2; side < 4; side++) {
        int bc = side == 2 ? g->bc_left : g->bc_right;
        if (bc != 1 && bc != 2) continue;
        orc_esat(side, cz, cx, bc == 2 ? px : NULL, nxt, c->f_lr_f1, c->f_lr_f2, c->f_imp_px, nx, nz, dx, dz, dt);
        orc_esat_dn(side, cx, nxt, c->f_dxc, nx, nz, dx, dz);
        orc_esat(side, cx, cz, bc == 2 ? pz : NULL, nzt, mu, mu, c->f_imp_s, nx, nz, dx, dz, dt);
    }
}

/* Spatial right-hand side L(curr) = volume terms + SAT terms (no source, no time
   update): out = [2][nx][nz].  prev may be NULL (treated as zero). */
int orc_rhs(const orc_grid* g, const real* model, const real* curr, const real* prev, real* out)
{
    d2b_init();
    const long nxz = (long)g->nx * g->nz;
    if (g->nx < 12 || g->nz < 12) return 1;
    real* buf = (real*)calloc((size_t)nxz * 16, sizeof(real));
    if (!buf) return 3;
    coef_fields c;
    coef_init(g, model, buf, &c);
    real* ux = buf + 9 * nxz;
    real* uzd = buf + 11 * nxz;
    real* tmp = buf + 13 * nxz;
    real* zero = buf + 14 * nxz;  /* 2 fields of zeros */
    const real* px = prev ? prev : zero;
    const real* pz = prev ? prev + nxz : zero + nxz;
    rhs_volume(g, model, &c, curr, curr + nxz, out, out + nxz, ux, uzd, tmp);
    rhs_sat(g, model, &c, curr, curr + nxz, px, pz, out, out + nxz);
    free(buf);
    return 0;
}

/* The reference images with the adjoint derivatives of the previous adjoint step
   (we_op.cpp:680 runs before they are recomputed at 697-717).  Tests may switch this
   lag off (0) to check that the lambda and rho gradients are then the exact discrete
   gradient, or reverse it (-1: the derivatives of the next adjoint level, the adjoint
   value of the current one) to check that the lag's error changes sign with its
   direction; the default (1) is the reference behaviour. */
static int imaging_lag = 1;
void orc_set_imaging_lag(int lag) { imaging_lag = lag; }

static void adjoint_derivatives(const orc_grid* g, const real* curr, real* ux, real* uzd)
{
    const int nx = g->nx, nz = g->nz;
    const long nxz = (long)nx * nz;
    orc_Dx(0, curr, ux, NULL, nx, nz, g->dx, 1.0);
    orc_Dx(0, curr + nxz, ux + nxz, NULL, nx, nz, g->dx, 1.0);
    orc_Dz(0, curr, uzd, NULL, nx, nz, g->dz, 1.0);
    orc_Dz(0, curr + nxz, uzd + nxz, NULL, nx, nz, g->dz, 1.0);
}

/* inj*: injector tables and traces [ninj][nt] per component (sources in forward mode,
   receivers in adjoint mode).  ext*: extractor tables and output traces (forward mode only).
   full: snapshots (forward writes when sub>0, adjoint reads).  grad != NULL selects the
   adjoint/imaging mode.  Returns 0 on success. */
static int propagate_pts(const orc_grid* g, const real* model, const pts* inj, const real* ib0, const real* ib1,
                         const pts* ext, real* ob0, real* ob1, real* full, real* grad)
{
    d2b_init();
    const int nx = g->nx, nz = g->nz, nt = g->nt, sub = g->sub, vti = g->vti;
    const long nxz = (long)nx * nz;
    const double dx = g->dx, dz = g->dz, dt = g->dt;
    if (nx < 12 || nz < 12 || nt < 2) return 1;
    if (grad && (sub <= 0 || !full)) return 2;

    real* work = (real*)calloc((size_t)nxz * 26, sizeof(real));
    if (!work) return 3;
    real* prev = work;              /* prev, curr, next: 3 x [2][nxz] */
    real* curr = work + 2 * nxz;
    real* next = work + 4 * nxz;
    real* ux = work + 6 * nxz;      /* (Dx u_x, Dx u_z) */
    real* uzd = work + 8 * nxz;     /* (Dz u_x, Dz u_z) */
    real* tmp = work + 10 * nxz;    /* 1 field */
    real* gtmp = work + 11 * nxz;   /* 4 fields for the imaging condition */
    coef_fields c;
    coef_init(g, model, work + 15 * nxz, &c);   /* 9 fields */
    real* lead_curr = work + 24 * nxz;          /* 2 fields: adjoint level imaged after the step (lag -1) */
    const real* rho = model + 2 * nxz;

    /* initial condition: prev = 1/2 dt^2/rho * inject(src[0])  (we_op.cpp:655-663) */
    pts_inject(inj, ib0, ib1, 0, prev, prev + nxz);
    for (long i = 0; i < nxz; i++) {
        prev[i] *= (real)(0.5 * dt * dt) / rho[i];
        prev[nxz + i] *= (real)(0.5 * dt * dt) / rho[i];
    }

    for (int it = 0; it < nt - 1; it++) {
        if (!grad && sub > 0 && it % sub == 0 && full)
            memcpy(full + (long)(it / sub) * 2 * nxz, curr, 2 * nxz * sizeof(real));
        if (!grad && ext) pts_extract(ext, ob0, ob1, it, curr, curr + nxz);
        const int image_now = grad && (nt - 1 - it) % sub == 0 && it != 0;
        if (image_now && imaging_lag >= 0) {
            if (!imaging_lag) adjoint_derivatives(g, curr, ux, uzd);
            compute_gradients(g, model, full, curr, ux, uzd, gtmp, grad, (nt - 1 - it) / sub, sub * dt);
        }
        if (image_now && imaging_lag < 0) memcpy(lead_curr, curr, 2 * nxz * sizeof(real));

        real* cx = curr;
        real* cz = curr + nxz;
        real* nxt = next;
        real* nzt = next + nxz;
        rhs_volume(g, model, &c, cx, cz, nxt, nzt, ux, uzd, tmp);
        /* sources (we_op.cpp:722-724) */
        pts_inject(inj, ib0, ib1, it, nxt, nzt);
        const real* px = prev;
        const real* pz = prev + nxz;
        rhs_sat(g, model, &c, cx, cz, px, pz, nxt, nzt);

        /* leapfrog (we_op.cpp:824-830) */
        const real dt2 = (real)(dt * dt);
        for (long i = 0; i < nxz; i++) {
            nxt[i] = dt2 * nxt[i] / rho[i] + 2 * cx[i] - px[i];
            nzt[i] = dt2 * nzt[i] / rho[i] + 2 * cz[i] - pz[i];
        }
        orc_scale_boundaries(nxt, nzt, model, vti, nx, nz, dx, dz, dt,
                             g->bc_top == 2, g->bc_bottom == 2, g->bc_left == 2, g->bc_right == 2);
        /* taper curr and next (we_op.cpp:846-861) */
        orc_taper(cx, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        orc_taper(cz, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        orc_taper(nxt, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        orc_taper(nzt, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        if (image_now && imaging_lag < 0) {
            adjoint_derivatives(g, next, ux, uzd);
            compute_gradients(g, model, full, lead_curr, ux, uzd, gtmp, grad, (nt - 1 - it) / sub, sub * dt);
        }

        real* b = prev; prev = curr; curr = next; next = b;
    }

    if (!grad && sub > 0 && (nt - 1) % sub == 0 && full)
        memcpy(full + (long)((nt - 1) / sub) * 2 * nxz, curr, 2 * nxz * sizeof(real));
    if (!grad && ext) pts_extract(ext, ob0, ob1, nt - 1, curr, curr + nxz);
    if (grad) {
        if (imaging_lag <= 0) adjoint_derivatives(g, curr, ux, uzd);
        compute_gradients(g, model, full, curr, ux, uzd, gtmp, grad, 0, sub * dt);
    }
    free(work);
    return 0;
}

/* propagate with delta_m3 injectors and extractors (the point-force, point-receiver case):
   inx/inz [ninj][nt] source traces per component, outx/outz [next][nt] receiver traces
   (accumulated), full [nsnap][2][nx][nz] snapshots (forward) or read (adjoint with grad). */
int orc_propagate(const orc_grid* g, const real* model,
                  int ninj, const int* ixi, const int* izi, const real* ixw, const real* izw,
                  const real* inx, const real* inz,
                  int next_, const int* exi, const int* ezi, const real* exw, const real* ezw,
                  real* outx, real* outz, real* full, real* grad)
{
    pts inj, ext;
    pts_delta(&inj, ninj, ixi, izi, ixw, izw, 0, g->nt, g->nz);
    if (next_ > 0) pts_delta(&ext, next_, exi, ezi, exw, ezw, 1, g->nt, g->nz);
    const int rc = propagate_pts(g, model, &inj, inx, inz, next_ > 0 ? &ext : NULL, outx, outz, full, grad);
    pts_free(&inj);
    if (next_ > 0) pts_free(&ext);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Acoustic operator nl_we_op_a (we_op.cpp:1863-2135): one pressure-like field, */
/* model (K, 1/rho) = (rho vp^2, 1/rho)                                         */
/* ------------------------------------------------------------------------- */
/* cell k rows (or columns) into side `side` (0 top, 1 bottom, 2 left, 3 right) on line l */
static inline long aside_cell(int side, int l, int k, int nx, int nz)
{
    if (side == 0) return (long)l * nz + k;
    if (side == 1) return (long)l * nz + nz - 1 - k;
    if (side == 2) return (long)k * nz + l;
    return (long)(nx - 1 - k) * nz + l;
}

/* asat_dirichlet_*, asat_absorbing_*, asat_neumann_* (spatial_operators.cpp:551-777),
   added into out over the whole side: kind 1 Dirichlet (the free surface; a its
   stiffness parameter), 2 absorbing (in1 the previous level), 3 Neumann. */
void orc_asat(int side, int kind, const real* in0, const real* in1, real* out, const real* K, const real* ri,
              int nx, int nz, double dx, double dz, double dt, double a)
{
    const int nline = side < 2 ? nx : nz;
    const double d = side < 2 ? dz : dx;
    for (int l = 0; l < nline; l++) {
        const long b = aside_cell(side, l, 0, nx, nz);
        if (kind == 1) {
            /* + H^-1 S^T (ri in0)_0, - H^-1 (ri in0 / (h a))_0 */
            for (int k = 0; k < 4; k++)
                out[aside_cell(side, l, k, nx, nz)] += (real)(1.0 / (HCOEF[k] * d * d)) * (real)SCOEF[k] * ri[b] * in0[b];
            out[b] -= ri[b] * in0[b] / (real)(a * HCOEF[0] * d * d);
        } else {
            /* - H^-1 (ri S in0 [- in1 / (2 dt rho v)])_0 */
            real sum = 0;
            for (int k = 0; k < 4; k++) sum += (real)SCOEF[k] * in0[aside_cell(side, l, k, nx, nz)];
            real v = ri[b] * sum / (real)d;
            if (kind == 2) v -= in1[b] / ((real)(2 * dt) * (real)sqrt(K[b] / ri[b]));
            out[b] -= (real)(1.0 / (HCOEF[0] * d)) * v;
        }
    }
}

/* asat_scale_boundaries (spatial_operators.cpp:779-860): a cell on an absorbing side is
   divided by 1 + the sum, over the absorbing sides it lies on, of v dt / (2 h H0) with
   v = sqrt(K / rho) of the cell (the corners take both of their sides, as the reference's
   sc1/sc2 bookkeeping does). */
void orc_asat_scale(real* u, const real* K, const real* ri, int nx, int nz, double dx, double dz, double dt,
                    int top, int bottom, int left, int right)
{
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            const int on_t = top && iz == 0, on_b = bottom && iz == nz - 1;
            const int on_l = left && ix == 0, on_r = right && ix == nx - 1;
            if (!(on_t || on_b || on_l || on_r)) continue;
            const long i = (long)ix * nz + iz;
            const real v = (real)sqrt(K[i] * ri[i]) * (real)dt;
            real s = 0;
            if (on_t) s += v / (real)(2 * dz * HCOEF[0]);
            if (on_l) s += v / (real)(2 * dx * HCOEF[0]);
            if (on_b) s += v / (real)(2 * dz * HCOEF[0]);
            if (on_r) s += v / (real)(2 * dx * HCOEF[0]);
            u[i] = u[i] / (1 + s);
        }
}

/* nl_we_op_a::compute_gradients (we_op.cpp:1883-1935): K gradient (with the -1/K^2 of
   the K^-1 parameterisation) from the second time difference of the snapshots against
   the adjoint level, 1/rho gradient from first derivatives; grad [2][nx][nz] accumulated */
static void acoustic_gradients(const orc_grid* g, const real* model, const real* full, const real* adj, real* tmp,
                               real* grad, int idx, double dtg)
{
    const int nx = g->nx, nz = g->nz;
    const long nxz = (long)nx * nz;
    const real* K = model;
    const real* F0 = full + (long)idx * nxz;
    real *t0 = tmp, *t1 = tmp + nxz, *t2 = tmp + 2 * nxz, *t3 = tmp + 3 * nxz;
    orc_Dx(0, F0, t0, NULL, nx, nz, g->dx, 1.0);
    orc_Dz(0, F0, t1, NULL, nx, nz, g->dz, 1.0);
    orc_Dx(0, adj, t2, NULL, nx, nz, g->dx, 1.0);
    orc_Dz(0, adj, t3, NULL, nx, nz, g->dz, 1.0);
    const double t = idx * dtg;
    double w = 1;
    if (g->gdamp_tmax - t > 0 && t - g->gdamp_tmin >= 0)
        w = cos(g->gdamp_power * M_PI * (t - g->gdamp_tmin) / (g->gdamp_tmax - g->gdamp_tmin));
    w *= w;
    const int last = g->nt / g->sub;
    for (long i = 0; i < nxz; i++) {
        real d2;
        real wr = (real)(w * dtg);
        if (idx > 0 && idx < last) {
            d2 = full[(long)(idx + 1) * nxz + i] - 2 * F0[i] + full[(long)(idx - 1) * nxz + i];
        } else if (idx == 0) {
            d2 = full[(long)(idx + 1) * nxz + i] - F0[i];
            wr = (real)(w * 0.5 * dtg);
        } else {
            d2 = -F0[i] + full[(long)(idx - 1) * nxz + i];
        }
        grad[i] -= (real)w / ((real)dtg * K[i] * K[i]) * adj[i] * d2;
        grad[nxz + i] += wr * (t2[i] * t0[i] + t3[i] * t1[i]);
    }
}

/* nl_we_op_a::propagate (we_op.cpp:1937-2135) with point injectors and extractors of one
   component; the boundary kinds come from bc_* (1 Dirichlet free surface, 2 absorbing,
   3 Neumann), alpha is the Dirichlet SAT parameter 0.2508560249 / free_surface_stiffness */
static int propagate_acoustic_pts(const orc_grid* g, const real* model, double alpha, const pts* inj, const real* ib,
                                  const pts* ext, real* ob, real* full, real* grad)
{
    d2b_init();
    const int nx = g->nx, nz = g->nz, nt = g->nt, sub = g->sub;
    const long nxz = (long)nx * nz;
    const double dx = g->dx, dz = g->dz, dt = g->dt;
    if (nx < 12 || nz < 12 || nt < 2) return 1;
    if (grad && (sub <= 0 || !full)) return 2;
    const int bc[4] = {g->bc_top, g->bc_bottom, g->bc_left, g->bc_right};
    for (int k = 0; k < 4; k++)
        if (bc[k] < 1 || bc[k] > 3) return 4;
    real* work = (real*)calloc((size_t)nxz * 7, sizeof(real));
    if (!work) return 3;
    real* prev = work;
    real* curr = work + nxz;
    real* next = work + 2 * nxz;
    real* tmp = work + 3 * nxz;  /* 4 fields for the gradients */
    const real* K = model;
    const real* ri = model + nxz;

    /* prev = K/2 dt^2 inject(src[0])  (we_op.cpp:1977-1979) */
    pts_inject(inj, ib, NULL, 0, prev, NULL);
    for (long i = 0; i < nxz; i++) prev[i] *= (real)0.5 * K[i] * (real)dt * (real)dt;

    for (int it = 0; it < nt - 1; it++) {
        if (!grad && sub > 0 && it % sub == 0 && full) memcpy(full + (long)(it / sub) * nxz, curr, nxz * sizeof(real));
        if (!grad && ext) pts_extract(ext, ob, NULL, it, curr, NULL);
        if (grad && (nt - 1 - it) % sub == 0 && it != 0)
            acoustic_gradients(g, model, full, curr, tmp, grad, (nt - 1 - it) / sub, sub * dt);
        orc_Dxx_var(0, curr, next, ri, nx, nz, dx, 1.0);
        orc_Dzz_var(1, curr, next, ri, nx, nz, dz, 1.0);
        pts_inject(inj, ib, NULL, it, next, NULL);
        for (int side = 0; side < 4; side++)
            orc_asat(side, bc[side], curr, prev, next, K, ri, nx, nz, dx, dz, dt, alpha);
        for (long i = 0; i < nxz; i++) next[i] = (real)dt * (real)dt * next[i] * K[i] + 2 * curr[i] - prev[i];
        orc_asat_scale(next, K, ri, nx, nz, dx, dz, dt, bc[0] == 2, bc[1] == 2, bc[2] == 2, bc[3] == 2);
        orc_taper(curr, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        orc_taper(next, nx, nz, g->taper_top, g->taper_bottom, g->taper_left, g->taper_right, g->taper_strength);
        real* b = prev; prev = curr; curr = next; next = b;
    }
    if (!grad && sub > 0 && (nt - 1) % sub == 0 && full) memcpy(full + (long)((nt - 1) / sub) * nxz, curr, nxz * sizeof(real));
    if (!grad && ext) pts_extract(ext, ob, NULL, nt - 1, curr, NULL);
    if (grad) acoustic_gradients(g, model, full, curr, tmp, grad, 0, sub * dt);
    free(work);
    return 0;
}

/* delta_m3 footprints of one component (the acoustic pressure) */
static void pts_delta1(pts* p, int ntr, const int* xi, const int* zi, const real* xw, const real* zw, int row, int nt,
                       int nz)
{
    pts_alloc(p, 9 * ntr);
    for (int t = 0; t < ntr; t++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                pts_add(p, 0, (long)xi[3 * t + a] * nz + zi[3 * t + b], (long)t * nt,
                        xw[6 * t + 3 * row + a] * zw[6 * t + 3 * row + b]);
}

/* acoustic propagate with delta_m3 injectors and extractors: in [ninj][nt] source traces,
   out [next][nt] receiver traces (accumulated), full [nsnap][nx][nz] snapshots (forward)
   or read (adjoint with grad [2][nx][nz]); model (K, 1/rho) */
int orc_propagate_acoustic(const orc_grid* g, const real* model, double alpha, int ninj, const int* ixi,
                           const int* izi, const real* ixw, const real* izw, const real* in, int next_,
                           const int* exi, const int* ezi, const real* exw, const real* ezw, real* out, real* full,
                           real* grad)
{
    pts inj, ext;
    pts_delta1(&inj, ninj, ixi, izi, ixw, izw, 0, g->nt, g->nz);
    if (next_ > 0) pts_delta1(&ext, next_, exi, ezi, exw, ezw, 1, g->nt, g->nz);
    const int rc = propagate_acoustic_pts(g, model, alpha, &inj, in, next_ > 0 ? &ext : NULL, out, full, grad);
    pts_free(&inj);
    if (next_ > 0) pts_free(&ext);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Model preparation: analyzeModel (we_op.cpp:267-474) and convert_model         */
/* (we_op.cpp:476-494, 1441-1462)                                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    double vpmin, vpmax, vsmin, vsmax, rhomin, rhomax, deltamin, deltamax, epsmin, epsmax;
    int soft_clip;
    double courant, dt;    /* requested par.courant, par.dt (<0 auto, 0 courant-based, >0 capped) */
    int sub;               /* requested par.sub (<0 auto) */
    int nt_data; double dt_data;
    /* outputs */
    double vmax, vmin, dt_prop, courant_out;
    int nt_prop, sub_out;
} orc_model_par;

/* model: [3|5][nx][nz] (vp, vs, rho[, delta, eps]) modified in place. */
void orc_analyze_model(orc_model_par* p, real* model, int nparam, int nx, int nz, double dx, double dz)
{
    const long nxz = (long)nx * nz;
    real* vp = model;
    real* vs = model + nxz;
    real* rho = model + 2 * nxz;
#define CLIP(a, lo, hi) do { for (long i = 0; i < nxz; i++) { real v_ = (a)[i]; if (v_ > (real)(hi)) v_ = (real)(hi); if (v_ < (real)(lo)) v_ = (real)(lo); (a)[i] = v_; } } while (0)
    if (nparam == 2) {  /* acoustic (vp, rho): always hard-clipped, vmin/vmax from vp (we_op.cpp:401-428) */
        CLIP(vp, p->vpmin, p->vpmax);
        CLIP(model + nxz, p->rhomin, p->rhomax);
    }
    if (nparam == 5) {
        real* del = model + 3 * nxz;
        for (long i = 0; i < nxz; i++) {
            real v1 = (real)(-0.5 * (1 - vs[i] * vs[i] / (vp[i] * vp[i])));
            real v2 = (real)(2.0 / (vp[i] * vp[i] / (vs[i] * vs[i]) - 1));
            if (del[i] < v1) del[i] = (real)(0.9999 * v1);
            else if (del[i] > v2) del[i] = (real)(0.9999 * v2);
        }
    }
    if (nparam != 2 && !p->soft_clip) {
        CLIP(vp, p->vpmin, p->vpmax);
        CLIP(vs, p->vsmin, p->vsmax);
        CLIP(rho, p->rhomin, p->rhomax);
        if (nparam == 5) {
            CLIP(model + 3 * nxz, p->deltamin, p->deltamax);
            CLIP(model + 4 * nxz, p->epsmin, p->epsmax);
        }
        for (long i = 0; i < nxz; i++)
            if (vp[i] < (real)sqrt(2.0) * vs[i]) vs[i] = (real)(vp[i] / sqrt(2.0001));
    }
#undef CLIP
    double vpmx = -1e300, vsmn = 1e300;
    for (long i = 0; i < nxz; i++) {
        if (vp[i] > vpmx) vpmx = vp[i];
        const real vmin_i = nparam == 2 ? vp[i] : vs[i];
        if (vmin_i < vsmn) vsmn = vmin_i;
    }
    p->vmax = vpmx;
    p->vmin = vsmn;
    if (nparam == 5) {
        const real* e = model + 4 * nxz;
        for (long i = 0; i < nxz; i++) {
            double v = (double)(real)(vp[i] * (real)sqrt(1 + 2 * e[i]));
            if (v > p->vmax) p->vmax = v;
        }
    }
    double hmin = dx < dz ? dx : dz;
    double tmax = (p->nt_data - 1) * p->dt_data;
    double dt = p->dt, courant = p->courant;
    if (dt == 0) dt = courant * hmin / p->vmax;
    else if (dt > 0) { dt = dt < p->dt_data ? dt : p->dt_data; courant = dt * p->vmax / hmin; }
    else {
        dt = courant * hmin / p->vmax;
        if (p->dt_data > dt) dt = p->dt_data / ceil(p->dt_data / dt);
        else dt = p->dt_data;
        courant = dt * p->vmax / hmin;
    }
    p->dt_prop = dt;
    p->courant_out = courant;
    p->nt_prop = (int)lround(tmax / dt) + 1;
    p->sub_out = p->sub < 0 ? (int)lround(p->dt_data / dt) : p->sub;
}

/* forward: (vp, vs, rho[, delta, eps]) -> (lambda, mu, rho[, c13, eps]) */
void orc_convert_model(real* m, int nparam, long n)
{
    if (nparam == 2) {  /* nl_we_op_a::convert_model (we_op.cpp:1863-1881): (vp, rho) -> (K, 1/rho) */
        for (long i = 0; i < n; i++) {
            m[i] = m[n + i] * m[i] * m[i];
            m[n + i] = (real)1 / m[n + i];
        }
        return;
    }
    for (long i = 0; i < n; i++) {
        real vp = m[i], vs = m[n + i], rho = m[2 * n + i];
        m[i] = rho * (vp * vp - 2 * vs * vs);
        m[n + i] = rho * vs * vs;
        if (nparam == 5) {
            real l = m[i], mu = m[n + i];
            m[3 * n + i] = (real)sqrt(2 * (l + 2 * mu) * (l + mu) * m[3 * n + i] + (l + mu) * (l + mu)) - mu;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Trace processing                                                            */
/* ------------------------------------------------------------------------- */
static inline double sincf(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

/* sinc_resampler::apply_forward / apply_adjoint (operator.cpp:808-856), alpha = 0.5.
   Domain axis (ntm, dtm), range axis (ntd, dtd), both with origin 0; ntr traces. */
void orc_sinc(int adjoint, int add, int ntr, int ntm, double dtm, int ntd, double dtd, int hl,
              real* mod, real* dat)
{
    const double a = 0.5;
    if (adjoint && !add) memset(mod, 0, (size_t)ntr * ntm * sizeof(real));
    for (int x = 0; x < ntr; x++)
        for (int itd = 0; itd < ntd; itd++) {
            real t = (real)(itd * dtd);
            int itm0 = (int)floor((double)t / dtm);
            int lo = itm0 - hl + 1 > 0 ? itm0 - hl + 1 : 0;
            int hi = itm0 + hl < ntm - 1 ? itm0 + hl : ntm - 1;
            if (!adjoint) {
                real v = add ? dat[(long)x * ntd + itd] : (real)0;
                for (int itm = lo; itm <= hi; itm++)
                    v += mod[(long)x * ntm + itm] * (real)(sincf(M_PI * ((double)t - itm * (double)(real)dtm) / dtm) *
                                                           (a + (1 - a) * cos(M_PI * (itm - (double)t / dtm) / hl)));
                dat[(long)x * ntd + itd] = v;
            } else {
                for (int itm = lo; itm <= hi; itm++)
                    mod[(long)x * ntm + itm] += dat[(long)x * ntd + itd] *
                        (real)(sincf(M_PI * ((double)t - itm * (double)(real)dtm) / dtm) * (a + (1 - a) * cos(M_PI * (itm - (double)t / dtm) / hl)));
            }
        }
}

/* linear_resampler (operator.cpp:760-806) */
void orc_linear(int adjoint, int add, int ntr, int ntm, double dtm, int ntd, double dtd, real* mod, real* dat)
{
    if (adjoint && !add) memset(mod, 0, (size_t)ntr * ntm * sizeof(real));
    for (int x = 0; x < ntr; x++)
        for (int it = 0; it < ntd; it++) {
            int i = (int)floor(it * dtd / dtm); if (i > ntm - 1) i = ntm - 1;
            int j = (int)ceil(it * dtd / dtm); if (j > ntm - 1) j = ntm - 1;
            real wj = (real)((it * dtd - i * dtm) / dtm), wi = 1 - wj;
            if (!adjoint) dat[(long)x * ntd + it] = (add ? dat[(long)x * ntd + it] : 0) + wi * mod[(long)x * ntm + i] + wj * mod[(long)x * ntm + j];
            else { mod[(long)x * ntm + i] += wi * dat[(long)x * ntd + it]; mod[(long)x * ntm + j] += wj * dat[(long)x * ntd + it]; }
        }
}

/* Dt forward / "adjoint" (misc.cpp:71-95); out must not alias in */
void orc_Dt(int adj, int ntr, int nt, double dt, const real* in, real* out)
{
    for (int x = 0; x < ntr; x++) {
        const real* a = in + (long)x * nt;
        real* o = out + (long)x * nt;
        if (!adj) {
            o[0] = (a[1] - a[0]) / (real)dt;
            for (int t = 1; t < nt - 1; t++) o[t] = (a[t + 1] - a[t - 1]) / (real)(2 * dt);
            o[nt - 1] = (a[nt - 1] - a[nt - 2]) / (real)dt;
        } else {
            o[0] = (a[1] + a[0]) / (real)dt;
            for (int t = 1; t < nt - 1; t++) o[t] = (a[t + 1] - a[t - 1]) / (real)(2 * dt);
            o[nt - 1] = (-a[nt - 1] - a[nt - 2]) / (real)dt;
        }
    }
}

/* applyHt (misc.cpp:45-69), in place */
void orc_Ht(int inv, int ntr, int nt, double dt, real* u)
{
    for (int x = 0; x < ntr; x++) {
        real* a = u + (long)x * nt;
        for (int t = 0; t < nt; t++) {
            double w = (t == 0 || t == nt - 1) ? 0.5 * dt : dt;
            a[t] = inv ? a[t] / (real)w : a[t] * (real)w;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Operator-level entry points: apply_forward (we_op.cpp:904-1055) and           */
/* apply_jacobianT (we_op.cpp:1057-1239) for point sources (mt=0) and point      */
/* receivers (gl=0), single process.                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    orc_grid g;            /* nx,nz,dx,dz,ox,oz, bc, taper, vti, gdamp; nt/dt/sub are outputs */
    orc_model_par mp;      /* bounds and time-axis requests */
    int ns;
    const double* sxz;     /* [ns][2] (x, z) */
    const int* nrec;       /* [ns] */
    const double* rxz;     /* [nr_total][2] */
    int seismotype;        /* 0 displacement, 1 velocity */
    int resampling;        /* 0 sinc, 1 linear */
    int sinc_hl;
    int mt;                /* 1: moment-tensor sources, allsrc [3][ns][nt] (mxx, mzz, mxz) */
    double gl;             /* > 0: DAS receivers of this gauge length (dipole_m3) */
    const double* rdip;    /* [nr_total] receiver dips (rad), used when gl > 0 */
    int acoustic;          /* 1: nl_we_op_a, model (vp, rho), allsrc [1][ns][nt], data [1][nr][nt] */
    double stiffness;      /* acoustic free-surface SAT stiffness (>= 1 keeps the exact limit) */
} orc_op;

static void op_prepare(orc_op* op, const real* pmod, real** model_out)
{
    const int np = op->acoustic ? 2 : (op->g.vti ? 5 : 3);
    const long nxz = (long)op->g.nx * op->g.nz;
    real* m = (real*)malloc((size_t)np * nxz * sizeof(real));
    memcpy(m, pmod, (size_t)np * nxz * sizeof(real));
    orc_analyze_model(&op->mp, m, np, op->g.nx, op->g.nz, op->g.dx, op->g.dz);
    op->g.nt = op->mp.nt_prop;
    op->g.dt = op->mp.dt_prop;
    op->g.sub = op->mp.sub_out;
    orc_convert_model(m, np, nxz);
    *model_out = m;
}

static int nr_total(const orc_op* op)
{
    int n = 0;
    for (int s = 0; s < op->ns; s++) n += op->nrec[s];
    return n;
}

static void resample(const orc_op* op, int adjoint, int add, int ntr, int ntm, double dtm, int ntd, double dtd, real* mod, real* dat)
{
    if (op->resampling == 1) orc_linear(adjoint, add, ntr, ntm, dtm, ntd, dtd, mod, dat);
    else orc_sinc(adjoint, add, ntr, ntm, dtm, ntd, dtd, op->sinc_hl, mod, dat);
}

/* The receivers of shot s (first receiver r0, n of them) as an injector (row 0) or
   extractor (row 1): delta_m3 points, or dipole_m3 gauges when op->gl > 0; traces at
   offsets r * nt from the component bases. */
static void receiver_pts(const orc_op* op, int r0, int n, int row, int nt, pts* p)
{
    const int nx = op->g.nx, nz = op->g.nz;
    double* rx = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* rz = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int r = 0; r < n; r++) { rx[r] = op->rxz[2 * (r0 + r)]; rz[r] = op->rxz[2 * (r0 + r) + 1]; }
    const int w = op->gl > 0 ? 6 : 3;
    int* xi = (int*)malloc(sizeof(int) * w * (n > 0 ? n : 1));
    int* zi = (int*)malloc(sizeof(int) * w * (n > 0 ? n : 1));
    real* xw = (real*)malloc(sizeof(real) * 2 * w * (n > 0 ? n : 1));
    real* zw = (real*)malloc(sizeof(real) * 2 * w * (n > 0 ? n : 1));
    if (op->gl > 0) {
        orc_dipole_m3(nx, nz, op->g.dx, op->g.dz, op->g.ox, op->g.oz, n, rx, rz, op->rdip + r0, op->gl, xi, zi, xw, zw);
        pts_dipole(p, n, xi, zi, xw, zw, row, nt, nz);
    } else {
        orc_delta_m3(nx, nz, op->g.dx, op->g.dz, op->g.ox, op->g.oz, n, rx, rz, xi, zi, xw, zw);
        pts_delta(p, n, xi, zi, xw, zw, row, nt, nz);
    }
    free(rx); free(rz); free(xi); free(zi); free(xw); free(zw);
}

/* The acoustic operator (nl_we_op_a through nl_we_op_e::apply_forward / apply_jacobianT,
   we_op.cpp:904-1239 with nmodels = 2): one source and one data component, delta_m3
   points, propagate_acoustic_pts per shot, snapshots [ns][1+nt/sub][nx][nz]. */
static double acoustic_alpha(const orc_op* op)
{
    return 0.2508560249 / (op->stiffness > 1.0 ? op->stiffness : 1.0);
}

static void acoustic_pts(const orc_op* op, int n, const double* xz, int row, int nt, pts* p)
{
    const int nx = op->g.nx, nz = op->g.nz, m = n > 0 ? n : 1;
    double* x = (double*)malloc(sizeof(double) * m);
    double* z = (double*)malloc(sizeof(double) * m);
    for (int r = 0; r < n; r++) { x[r] = xz[2 * r]; z[r] = xz[2 * r + 1]; }
    int* xi = (int*)malloc(sizeof(int) * 3 * m);
    int* zi = (int*)malloc(sizeof(int) * 3 * m);
    real* xw = (real*)malloc(sizeof(real) * 6 * m);
    real* zw = (real*)malloc(sizeof(real) * 6 * m);
    orc_delta_m3(nx, nz, op->g.dx, op->g.dz, op->g.ox, op->g.oz, n, x, z, xi, zi, xw, zw);
    pts_delta1(p, n, xi, zi, xw, zw, row, nt, nz);
    free(x); free(z); free(xi); free(zi); free(xw); free(zw);
}

static int apply_forward_a(orc_op* op, int add, const real* m, const real* allsrc, real* data, real* full)
{
    const int nz = op->g.nz, nt = op->g.nt, ns = op->ns, ntd = op->mp.nt_data;
    const double dtd = op->mp.dt_data, dt = op->g.dt;
    const long nxz = (long)op->g.nx * nz;
    const int nr = nr_total(op);
    real* src = (real*)calloc((size_t)ns * nt, sizeof(real));
    real* rcv = (real*)calloc((size_t)(nr > 0 ? nr : 1) * nt, sizeof(real));
    resample(op, 0, 0, ns, ntd, dtd, nt, dt, (real*)allsrc, src);
    const int nsnap = op->g.sub > 0 ? 1 + nt / op->g.sub : 0;
    int r0 = 0, rc = 0;
    for (int s = 0; s < ns && !rc; s++) {
        pts inj, ext;
        acoustic_pts(op, 1, op->sxz + 2 * s, 0, nt, &inj);
        const int n = op->nrec[s];
        if (n > 0) acoustic_pts(op, n, op->rxz + 2 * (long)r0, 1, nt, &ext);
        real* fs = full && nsnap ? full + (long)s * nsnap * nxz : NULL;
        rc = propagate_acoustic_pts(&op->g, m, acoustic_alpha(op), &inj, src + (long)s * nt, n > 0 ? &ext : NULL,
                                    rcv + (long)r0 * nt, fs, NULL);
        pts_free(&inj);
        if (n > 0) pts_free(&ext);
        r0 += n;
    }
    if (!rc) {
        if (op->seismotype == 1) {
            real* t = (real*)calloc((size_t)(nr > 0 ? nr : 1) * nt, sizeof(real));
            orc_Dt(0, nr, nt, dt, rcv, t);
            free(rcv);
            rcv = t;
        }
        real alpha = (real)(dt / dtd);
        for (long i = 0; i < (long)nr * nt; i++) rcv[i] *= alpha;
        resample(op, 1, add, nr, ntd, dtd, nt, dt, data, rcv);
    }
    free(src); free(rcv);
    return rc;
}

static int apply_jacobianT_a(orc_op* op, int add, real* grad, const real* m, const real* data, const real* full)
{
    const int nx = op->g.nx, nz = op->g.nz, nt = op->g.nt, ns = op->ns, ntd = op->mp.nt_data;
    const double dtd = op->mp.dt_data, dt = op->g.dt;
    const long nxz = (long)nx * nz;
    const int nr = nr_total(op);
    if (op->g.sub <= 0) return 2;
    const int nsnap = 1 + nt / op->g.sub;
    real* g = (real*)calloc((size_t)2 * nxz, sizeof(real));
    real* rcv = (real*)calloc((size_t)(nr > 0 ? nr : 1) * nt, sizeof(real));
    resample(op, 0, 0, nr, ntd, dtd, nt, dt, (real*)data, rcv);
    orc_Ht(1, nr, nt, dt, rcv);
    real alpha = (real)(dt / dtd);
    for (int x = 0; x < nr; x++) {
        real* a = rcv + (long)x * nt;
        for (int t = 0; t < nt / 2; t++) { real v = a[t]; a[t] = a[nt - 1 - t]; a[nt - 1 - t] = v; }
        for (int t = 0; t < nt; t++) a[t] *= -alpha;
    }
    if (op->seismotype == 1) {
        real* t = (real*)calloc((size_t)(nr > 0 ? nr : 1) * nt, sizeof(real));
        orc_Dt(1, nr, nt, dt, rcv, t);
        free(rcv);
        rcv = t;
    }
    int r0 = 0, rc = 0;
    for (int s = 0; s < ns && !rc; s++) {
        const int n = op->nrec[s];
        if (n > 0) {
            pts inj;
            acoustic_pts(op, n, op->rxz + 2 * (long)r0, 0, nt, &inj);
            rc = propagate_acoustic_pts(&op->g, m, acoustic_alpha(op), &inj, rcv + (long)r0 * nt, NULL, NULL,
                                        (real*)full + (long)s * nsnap * nxz, g);
            pts_free(&inj);
        }
        r0 += n;
    }
    if (!rc) {
        for (int c = 0; c < 2; c++) {
            orc_applyH(0, 0, g + c * nxz, nx, nz, op->g.dz);
            orc_applyH(1, 0, g + c * nxz, nx, nz, op->g.dx);
        }
        /* chain rule (we_op.cpp:1220-1229): g_vp = 2 sqrt(K/(1/rho)) g_K,
           g_rho = (1/rho) K g_K - (1/rho)^2 g_(1/rho) */
        for (long i = 0; i < nxz; i++) {
            const real K = m[i], ri = m[nxz + i], gk = g[i], gr = g[nxz + i];
            g[nxz + i] = ri * K * gk - ri * ri * gr;
            g[i] = (real)(2 * sqrt(K / ri)) * gk;
        }
        for (long i = 0; i < 2 * nxz; i++) grad[i] = (add ? grad[i] : 0) + g[i];
    }
    free(g); free(rcv);
    return rc;
}

/* allsrc: [2][ns][nt_data] point forces, or [3][ns][nt_data] moment tensors (mxx, mzz,
   mxz) when op->mt; data: [2][nr_total][nt_data]; full (optional, sub>0):
   [ns][1+nt/sub][2][nx][nz].  we_op.cpp:904-1055. */
int orc_apply_forward(orc_op* op, int add, const real* pmod, const real* allsrc, real* data, real* full)
{
    real* m;
    op_prepare(op, pmod, &m);
    if (op->acoustic) {
        const int rc = apply_forward_a(op, add, m, allsrc, data, full);
        free(m);
        return rc;
    }
    const int nx = op->g.nx, nz = op->g.nz, nt = op->g.nt, ns = op->ns, ntd = op->mp.nt_data;
    const double dtd = op->mp.dt_data, dt = op->g.dt;
    const long nxz = (long)nx * nz;
    const int nr = nr_total(op);
    const int nc = op->mt ? 3 : 2;
    real* src = (real*)calloc((size_t)nc * ns * nt, sizeof(real));
    real* rcv = (real*)calloc((size_t)2 * nr * nt, sizeof(real));
    resample(op, 0, 0, nc * ns, ntd, dtd, nt, dt, (real*)allsrc, src);
    const int nsnap = op->g.sub > 0 ? 1 + nt / op->g.sub : 0;
    int r0 = 0, rc = 0;
    for (int s = 0; s < ns && !rc; s++) {
        int ixi[3], izi[3];
        real ixw[12], izw[12];
        pts inj, ext;
        const real *ib0, *ib1;
        if (op->mt) {
            orc_ddelta_m3(nx, nz, op->g.dx, op->g.dz, op->g.ox, op->g.oz, 1, &op->sxz[2 * s], &op->sxz[2 * s + 1],
                          ixi, izi, ixw, izw);
            pts_ddelta(&inj, 1, ixi, izi, ixw, izw, (long)s * nt, (long)(ns + s) * nt, (long)(2 * ns + s) * nt, nt, nz);
            ib0 = ib1 = src;
        } else {
            orc_delta_m3(nx, nz, op->g.dx, op->g.dz, op->g.ox, op->g.oz, 1, &op->sxz[2 * s], &op->sxz[2 * s + 1],
                         ixi, izi, ixw, izw);
            pts_delta(&inj, 1, ixi, izi, ixw, izw, 0, nt, nz);
            ib0 = src + (long)s * nt;
            ib1 = src + (long)(ns + s) * nt;
        }
        const int n = op->nrec[s];
        receiver_pts(op, r0, n, 1, nt, &ext);
        real* fs = full && nsnap ? full + (long)s * nsnap * 2 * nxz : NULL;
        rc = propagate_pts(&op->g, m, &inj, ib0, ib1, n > 0 ? &ext : NULL, rcv + (long)r0 * nt,
                           rcv + (long)(nr + r0) * nt, fs, NULL);
        pts_free(&inj);
        pts_free(&ext);
        r0 += n;
    }
    if (!rc) {
        if (op->gl > 0) orc_dipole_to_strain(0, rcv, op->rdip, nr, nt, op->gl);
        if (op->seismotype == 1) {
            real* t = (real*)calloc((size_t)2 * nr * nt, sizeof(real));
            orc_Dt(0, 2 * nr, nt, dt, rcv, t);
            free(rcv);
            rcv = t;
        }
        real alpha = (real)(dt / dtd);
        for (long i = 0; i < (long)2 * nr * nt; i++) rcv[i] *= alpha;
        resample(op, 1, add, 2 * nr, ntd, dtd, nt, dt, data, rcv);
    }
    free(src); free(rcv); free(m);
    return rc;
}

/* Mathematical adjoint-state gradient: grad = add*grad + J^T data, with the reference's
   H quadrature and chain rule.  full: the snapshots written by orc_apply_forward on the
   same model.  grad: [3][nx][nz] (iso) or [5][nx][nz] (vti; components 3,4 set to 0). */
int orc_apply_jacobianT(orc_op* op, int add, real* grad, const real* pmod0, const real* data, const real* full)
{
    real* m;
    op_prepare(op, pmod0, &m);
    if (op->acoustic) {
        const int rc = apply_jacobianT_a(op, add, grad, m, data, full);
        free(m);
        return rc;
    }
    const int nx = op->g.nx, nz = op->g.nz, nt = op->g.nt, ns = op->ns, ntd = op->mp.nt_data;
    const double dtd = op->mp.dt_data, dt = op->g.dt;
    const long nxz = (long)nx * nz;
    const int np = op->g.vti ? 5 : 3;
    const int nr = nr_total(op);
    if (op->g.sub <= 0) { free(m); return 2; }
    const int nsnap = 1 + nt / op->g.sub;
    real* g = (real*)calloc((size_t)np * nxz, sizeof(real));
    real* rcv = (real*)calloc((size_t)2 * nr * nt, sizeof(real));
    resample(op, 0, 0, 2 * nr, ntd, dtd, nt, dt, (real*)data, rcv);
    orc_Ht(1, 2 * nr, nt, dt, rcv);
    /* reverse time, scale by -alpha */
    real alpha = (real)(dt / dtd);
    for (int x = 0; x < 2 * nr; x++) {
        real* a = rcv + (long)x * nt;
        for (int t = 0; t < nt / 2; t++) { real v = a[t]; a[t] = a[nt - 1 - t]; a[nt - 1 - t] = v; }
        for (int t = 0; t < nt; t++) a[t] *= -alpha;
    }
    if (op->seismotype == 1) {
        real* t = (real*)calloc((size_t)2 * nr * nt, sizeof(real));
        orc_Dt(1, 2 * nr, nt, dt, rcv, t);
        free(rcv);
        rcv = t;
    }
    if (op->gl > 0) orc_dipole_to_strain(1, rcv, op->rdip, nr, nt, op->gl);  /* we_op.cpp:1130-1135 */
    int r0 = 0, rc = 0;
    for (int s = 0; s < ns && !rc; s++) {
        const int n = op->nrec[s];
        pts inj;
        receiver_pts(op, r0, n, 0, nt, &inj);
        rc = propagate_pts(&op->g, m, &inj, rcv + (long)r0 * nt, rcv + (long)(nr + r0) * nt, NULL, NULL, NULL,
                           (real*)full + (long)s * nsnap * 2 * nxz, g);
        pts_free(&inj);
        r0 += n;
    }
    if (!rc) {
        for (int c = 0; c < 3; c++) {
            orc_applyH(0, 0, g + c * nxz, nx, nz, op->g.dz);
            orc_applyH(1, 0, g + c * nxz, nx, nz, op->g.dx);
        }
        /* chain rule (we_op.cpp:1208-1219) */
        const real* la = m;
        const real* mu = m + nxz;
        const real* rho = m + 2 * nxz;
        for (long i = 0; i < nxz; i++) {
            real gl = g[i], gm = g[nxz + i], gr = g[2 * nxz + i];
            g[2 * nxz + i] = (la[i] / rho[i]) * gl + (mu[i] / rho[i]) * gm + gr;
            g[nxz + i] = (real)(2 * sqrt(mu[i] * rho[i])) * (gm - 2 * gl);
            g[i] = (real)(2 * sqrt(rho[i] * (la[i] + 2 * mu[i]))) * gl;
        }
        for (long i = 0; i < (long)np * nxz; i++) grad[i] = (add ? grad[i] : 0) + g[i];
    }
    free(g); free(rcv); free(m);
    return rc;
}

/* Raw per-shot gradient (before H and chain rule) of one adjoint propagation,
   exposed for kernel-level parity tests: grad += imaging(adjoint traces). */
int orc_propagate_adjoint_raw(const orc_grid* g, const real* model, int n, const int* ixi, const int* izi,
                              const real* ixw, const real* izw, const real* inx, const real* inz,
                              const real* full, real* grad)
{
    return orc_propagate(g, model, n, ixi, izi, ixw, izw, inx, inz, 0, NULL, NULL, NULL, NULL, NULL, NULL, (real*)full, grad);
}
