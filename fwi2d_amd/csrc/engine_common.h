// Host-side pieces shared by the elastic and acoustic engines: device buffers, the
// per-cell lists of point injectors, and the delta_m3 point weights.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.h"
#include "sbp_tables.h"

namespace fwi2d {

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t err_ = (x);                                                                        \
        if (err_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(err_));               \
    } while (0)

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    void reserve(size_t b)
    {
        if (b <= bytes) return;
        release();
        HIPCHECK(hipMalloc(&p, b));
        bytes = b;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
    ~DevMem() { release(); }
};

// Cells touched by the point injectors of all shots, with their contributions (trace
// offset of the x component, weight) in injector order -- host build, device copy
struct InjList {
    std::vector<InjCell> cells;
    std::vector<int> ptr;
    std::vector<long long> ctx, ctz;  // trace offsets read by the x and z field components
    std::vector<float> cw;
    DevMem d_cells, d_ptr, d_ctx, d_ctz, d_cw;
    void build(const std::vector<PointEntry>& ents, hipStream_t s)
    {
        struct C {
            int shot, gx, gz, k;
            long long tx, tz;
            float w;
        };
        std::vector<C> all;
        all.reserve(ents.size() * 9);
        for (int k = 0; k < (int)ents.size(); k++) {
            const PointEntry& e = ents[k];
            for (int ix = 0; ix < 3; ix++)
                for (int iz = 0; iz < 3; iz++)
                    all.push_back({e.shot, e.xi0 + ix, e.zi0 + iz, k, e.trace, e.tz, e.wx[ix] * e.wz[iz]});
        }
        std::stable_sort(all.begin(), all.end(), [](const C& a, const C& b) {
            if (a.shot != b.shot) return a.shot < b.shot;
            if (a.gx != b.gx) return a.gx < b.gx;
            if (a.gz != b.gz) return a.gz < b.gz;
            return a.k < b.k;
        });
        cells.clear();
        ptr.assign(1, 0);
        ctx.clear();
        ctz.clear();
        cw.clear();
        for (size_t i = 0; i < all.size(); i++) {
            if (i == 0 || all[i].shot != all[i - 1].shot || all[i].gx != all[i - 1].gx || all[i].gz != all[i - 1].gz) {
                if (i > 0) ptr.push_back((int)ctx.size());
                cells.push_back({all[i].shot, all[i].gx, all[i].gz, 0});
            }
            ctx.push_back(all[i].tx);
            ctz.push_back(all[i].tz);
            cw.push_back(all[i].w);
        }
        if (!all.empty()) ptr.push_back((int)ctx.size());
        auto up = [&](DevMem& d, const void* h, size_t b) {
            d.reserve(std::max<size_t>(b, 16));
            if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, s));
        };
        up(d_cells, cells.data(), cells.size() * sizeof(InjCell));
        up(d_ptr, ptr.data(), ptr.size() * sizeof(int));
        up(d_ctx, ctx.data(), ctx.size() * sizeof(long long));
        up(d_ctz, ctz.data(), ctz.size() * sizeof(long long));
        up(d_cw, cw.data(), cw.size() * sizeof(float));
    }
};

// Footprint centre of a point: the nearest grid node, moved off the outermost rows
// (delta_m3, injector.cpp:3-64), and the point's offset from it in cells
struct Pivot {
    int px, pz;     // footprint centre
    double ax, az;  // position relative to it, in cells
};
inline Pivot point_pivot(int nx, int nz, double dx, double dz, double x, double z, double ox, double oz)
{
    int px = (int)std::lround((x - ox) / dx);
    int pz = (int)std::lround((z - oz) / dz);
    if (px == 0) px = 1;
    if (px == nx - 1) px = nx - 2;
    if (pz == 0) pz = 1;
    if (pz == nz - 1) pz = nz - 2;
    if (px < 1 || px > nx - 2 || pz < 1 || pz > nz - 2)
        throw std::runtime_error("source or receiver falls outside the model");
    return {px, pz, (x - dx * px - ox) / dx, (z - dz * pz - oz) / dz};
}
// injection weights (a^2 - a)/(2 h H), (1 - a^2)/(h H), (a^2 + a)/(2 h H) around pivot p;
// with deriv, the derivative weights (1 - 2a)/(2 h^2 H), 2a/(h^2 H), (-1 - 2a)/(2 h^2 H)
inline void inj_weights(double a, int p, int n, double h, bool deriv, float* w)
{
    const double h0 = h_weight(n, p - 1), h1 = h_weight(n, p), h2 = h_weight(n, p + 1);
    if (!deriv) {
        w[0] = (float)((a * a - a) / (2 * h * h0));
        w[1] = (float)((1 - a * a) / (h * h1));
        w[2] = (float)((a * a + a) / (2 * h * h2));
    } else {
        w[0] = (float)((1 - 2 * a) / (2 * h * h * h0));
        w[1] = (float)(2 * a / (h * h * h1));
        w[2] = (float)((-1 - 2 * a) / (2 * h * h * h2));
    }
}
// extraction (Lagrange) weights (a^2 - a)/2, 1 - a^2, (a^2 + a)/2
inline void ext_weights(double a, float* w)
{
    w[0] = (float)((a * a - a) / 2);
    w[1] = (float)(1 - a * a);
    w[2] = (float)((a * a + a) / 2);
}

// cos^2 taper factors of one axis of n cells with lo / hi taper widths (taperz/taperx,
// spatial_operators.cpp:523-549)
inline std::vector<float> taper_axis(int n, int lo, int hi, double strength)
{
    std::vector<float> f(n, 1.f);
    for (int i = 0; i < n; i++) {
        if (lo > 0 && i < lo) {
            float v = (float)std::cos(strength * 0.5 * M_PI * (double)(i + 1 - lo) / (double)(0 - lo));
            f[i] *= v * v;
        }
        if (hi > 0 && i >= n - hi) {
            float v = (float)std::cos(strength * 0.5 * M_PI * (double)(i - (n - hi)) / (double)hi);
            f[i] *= v * v;
        }
    }
    return f;
}

}  // namespace fwi2d
