// Test-only harness around the reference's own SBP sources, compiled from
// /root/reference/src as they lie (see oracle/Makefile, target `ref`).
//
// It exposes C entry points over the reference templates in
// include/spatial_operators.hpp (Dzz_var, Dxx_var, esat_neumann_*,
// esat_absorbing_*, esat_Dz_*, esat_Dx_*) and over the non-ISPC functions of
// cpp/spatial_operators.cpp (applyHz, applyHx, Dzz, Dxx, esat_scale_boundaries,
// vtisat_scale_boundaries).  The ISPC-backed functions of that file (Dx, Dz,
// mult_Dx, mult_Dz, taperx, taperz) are never referenced here, so the linker
// drops them and no stand-in for the ISPC objects is needed.
//
// Built with -DDOUBLE_PRECISION, so data_t is double.  Only the oracle tests load it.
#include "spatial_operators.hpp"

#define EXPORT extern "C" __attribute__((visibility("default")))

// coefficient "expressions": the test passes the evaluated coefficient arrays
static inline data_t p0(const data_t** par, int i) { return par[0][i]; }
static inline data_t p1(const data_t** par, int i) { return par[1][i]; }
static inline data_t p2(const data_t** par, int i) { return par[2][i]; }

EXPORT int ref_data_t_size() { return (int)sizeof(data_t); }

EXPORT void ref_Dzz_var(int add, const data_t* in, data_t* out, const data_t* c, int nx, int nz, data_t d, data_t a)
{
    const data_t* par[1] = {c};
    Dzz_var<p0>(add != 0, in, out, nx, nz, d, 0, nx, 0, nz, par, a);
}

EXPORT void ref_Dxx_var(int add, const data_t* in, data_t* out, const data_t* c, int nx, int nz, data_t d, data_t a)
{
    const data_t* par[1] = {c};
    Dxx_var<p0>(add != 0, in, out, nx, nz, d, 0, nx, 0, nz, par, a);
}

EXPORT void ref_Dzz(int add, const data_t* in, data_t* out, int nx, int nz, data_t d)
{
    Dzz(add != 0, in, out, nx, nz, d, 0, nx, 0, nz);
}

// side: 0 top, 1 bottom, 2 left, 3 right; in2/f3 == nullptr selects the free-surface
// (Neumann) SAT, otherwise the absorbing SAT.  out += SAT.
EXPORT void ref_esat(int side, const data_t* in0, const data_t* in1, const data_t* in2, data_t* out,
                     const data_t* f1, const data_t* f2, const data_t* f3, int nx, int nz, data_t dx, data_t dz, data_t dt)
{
    const data_t* par[3] = {f1, f2, f3};
    if (in2 == nullptr) {
        const data_t* in[2] = {in0, in1};
        switch (side) {
        case 0: esat_neumann_top<p0, p1>(true, in, out, nx, nz, dx, dz, 0, nx, par, 1.0); break;
        case 1: esat_neumann_bottom<p0, p1>(true, in, out, nx, nz, dx, dz, 0, nx, par, 1.0); break;
        case 2: esat_neumann_left<p0, p1>(true, in, out, nx, nz, dx, dz, 0, nz, par, 1.0); break;
        default: esat_neumann_right<p0, p1>(true, in, out, nx, nz, dx, dz, 0, nz, par, 1.0); break;
        }
    } else {
        const data_t* in[3] = {in0, in1, in2};
        switch (side) {
        case 0: esat_absorbing_top<p0, p1, p2>(true, in, out, nx, nz, dx, dz, dt, 0, nx, par, 1.0); break;
        case 1: esat_absorbing_bottom<p0, p1, p2>(true, in, out, nx, nz, dx, dz, dt, 0, nx, par, 1.0); break;
        case 2: esat_absorbing_left<p0, p1, p2>(true, in, out, nx, nz, dx, dz, dt, 0, nz, par, 1.0); break;
        default: esat_absorbing_right<p0, p1, p2>(true, in, out, nx, nz, dx, dz, dt, 0, nz, par, 1.0); break;
        }
    }
}

EXPORT void ref_esat_dn(int side, const data_t* in, data_t* out, const data_t* f, int nx, int nz, data_t dx, data_t dz)
{
    const data_t* par[1] = {f};
    switch (side) {
    case 0: esat_Dz_top<p0>(true, in, out, nx, nz, dz, 0, nx, par, 1.0); break;
    case 1: esat_Dz_bottom<p0>(true, in, out, nx, nz, dz, 0, nx, par, 1.0); break;
    case 2: esat_Dx_left<p0>(true, in, out, nx, nz, dx, 0, nz, par, 1.0); break;
    default: esat_Dx_right<p0>(true, in, out, nx, nz, dx, 0, nz, par, 1.0); break;
    }
}

// model: (lambda, mu, rho[, c13, eps]) as [np][nx][nz]
EXPORT void ref_scale_boundaries(data_t* ux, data_t* uz, const data_t* model, int vti, int nx, int nz,
                                 data_t dx, data_t dz, data_t dt, int top, int bottom, int left, int right)
{
    const long nxz = (long)nx * nz;
    const data_t* par[5] = {model, model + nxz, model + 2 * nxz, vti ? model + 3 * nxz : nullptr,
                            vti ? model + 4 * nxz : nullptr};
    data_t* in[2] = {ux, uz};
    if (vti) vtisat_scale_boundaries(in, nx, nz, dx, dz, 0, nx, 0, nz, par, dt, top, bottom, left, right);
    else esat_scale_boundaries(in, nx, nz, dx, dz, 0, nx, 0, nz, par, dt, top, bottom, left, right);
}

EXPORT void ref_applyH(int axis_x, int inv, const data_t* in, data_t* out, int nx, int nz, data_t d)
{
    if (axis_x) applyHx(inv != 0, false, in, out, nx, nz, d, 0, nx, 0, nz);
    else applyHz(inv != 0, false, in, out, nx, nz, d, 0, nx, 0, nz);
}

// Acoustic SAT terms (asat_dirichlet_*, asat_absorbing_*, asat_neumann_*) on one side over
// its whole length, out += SAT; kind 1 Dirichlet (stiffness parameter a), 2 absorbing
// (in1 the previous level), 3 Neumann; model (K, 1/rho).
EXPORT void ref_asat(int side, int kind, const data_t* in0, const data_t* in1, data_t* out, const data_t* K,
                     const data_t* ri, int nx, int nz, data_t dx, data_t dz, data_t dt, data_t a)
{
    const data_t* par[2] = {K, ri};
    const data_t* in[2] = {in0, in1};
    const int n = side < 2 ? nx : nz;
    if (kind == 1) {
        switch (side) {
        case 0: asat_dirichlet_top(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        case 1: asat_dirichlet_bottom(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        case 2: asat_dirichlet_left(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        default: asat_dirichlet_right(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        }
    } else if (kind == 2) {
        switch (side) {
        case 0: asat_absorbing_top(true, in, out, nx, nz, dx, dz, dt, 0, n, par, a); break;
        case 1: asat_absorbing_bottom(true, in, out, nx, nz, dx, dz, dt, 0, n, par, a); break;
        case 2: asat_absorbing_left(true, in, out, nx, nz, dx, dz, dt, 0, n, par, a); break;
        default: asat_absorbing_right(true, in, out, nx, nz, dx, dz, dt, 0, n, par, a); break;
        }
    } else {
        switch (side) {
        case 0: asat_neumann_top(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        case 1: asat_neumann_bottom(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        case 2: asat_neumann_left(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        default: asat_neumann_right(true, in, out, nx, nz, dx, dz, 0, n, par, a); break;
        }
    }
}

// asat_scale_boundaries over the whole grid; model (K, 1/rho)
EXPORT void ref_asat_scale(data_t* u, const data_t* K, const data_t* ri, int nx, int nz, data_t dx, data_t dz,
                           data_t dt, int top, int bottom, int left, int right)
{
    const data_t* par[2] = {K, ri};
    data_t* in[1] = {u};
    asat_scale_boundaries(in, nx, nz, dx, dz, 0, nx, 0, nz, par, dt, top, bottom, left, right);
}
