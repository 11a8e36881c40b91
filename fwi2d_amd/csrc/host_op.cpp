// Operator-level drop-in for the reference's nl_we_op_e / nl_we_op_vti
// (src/include/we_op.hpp:54-86, 150-176): the host pre/post-processing of
// apply_forward (src/cpp/we_op.cpp:904-1055) and apply_jacobianT
// (src/cpp/we_op.cpp:1057-1239) around the batched device engine.
//
// The per-trace time processing (resampling, Dt, Ht^-1, reversal, dt ratio) is composed
// here in double precision into one sparse matrix per direction, which the device applies
// to all traces with double accumulation; gradient post-processing (H quadrature, chain
// rule) runs here in double precision; the time stepping runs on the device in single
// precision.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "fwi2d_hip.h"
#include "trace_matrix.h"

namespace fwi2d {
extern thread_local std::string g_last_error;
}

namespace {

const double kPi = 3.14159265358979323846;

void check(int rc)
{
    if (rc) throw std::runtime_error(fwi2d_last_error());
}

// analyzeModel (we_op.cpp:267-474) for 3 or 5 parameters: bounds, the Vs <= Vp/sqrt(2)
// and Thomsen-delta admissibility clips, and the propagation time axis.
struct TimeAxis {
    int nt;
    double dt;
    int sub;
};

TimeAxis analyze_model(const fwi2d_op_desc& d, std::vector<double>& m)
{
    const long long n = (long long)d.nx * d.nz;
    double* vp = m.data();
    double* vs = vp + n;
    double* rho = vp + 2 * n;
    auto clip = [&](double* a, double lo, double hi) {
        for (long long i = 0; i < n; i++) a[i] = std::max(lo, std::min(hi, a[i]));
    };
    if (d.nparam == 2) {  // acoustic (vp, rho): always clipped (we_op.cpp:399-421)
        rho = vp + n;
        clip(vp, d.vpmin, d.vpmax);
        clip(rho, d.rhomin, d.rhomax);
    }
    if (d.nparam == 5) {
        double* del = vp + 3 * n;
        for (long long i = 0; i < n; i++) {
            const double v1 = -0.5 * (1 - vs[i] * vs[i] / (vp[i] * vp[i]));
            const double v2 = 2.0 / (vp[i] * vp[i] / (vs[i] * vs[i]) - 1);
            if (del[i] < v1) del[i] = 0.9999 * v1;
            else if (del[i] > v2) del[i] = 0.9999 * v2;
        }
    }
    if (d.nparam != 2 && !d.soft_clip) {
        clip(vp, d.vpmin, d.vpmax);
        clip(vs, d.vsmin, d.vsmax);
        clip(rho, d.rhomin, d.rhomax);
        if (d.nparam == 5) {
            clip(vp + 3 * n, d.deltamin, d.deltamax);
            clip(vp + 4 * n, d.epsilonmin, d.epsilonmax);
        }
        for (long long i = 0; i < n; i++)
            if (vp[i] < std::sqrt(2.0) * vs[i]) vs[i] = vp[i] / std::sqrt(2.0001);
    }
    double vmax = -1e300;
    for (long long i = 0; i < n; i++) vmax = std::max(vmax, vp[i]);
    if (d.nparam == 5)
        for (long long i = 0; i < n; i++) vmax = std::max(vmax, vp[i] * std::sqrt(1 + 2 * vp[4 * n + i]));
    const double hmin = std::min(d.dx, d.dz);
    const double tmax = (d.nt_data - 1) * d.dt_data;
    double dt = d.dt;
    if (dt == 0) dt = d.courant * hmin / vmax;
    else if (dt > 0) dt = std::min(dt, d.dt_data);
    else {
        dt = d.courant * hmin / vmax;
        dt = (d.dt_data > dt) ? d.dt_data / std::ceil(d.dt_data / dt) : d.dt_data;
    }
    TimeAxis t;
    t.dt = dt;
    t.nt = (int)std::lround(tmax / dt) + 1;
    t.sub = d.sub < 0 ? (int)std::lround(d.dt_data / dt) : d.sub;
    return t;
}

// convert_model (we_op.cpp:476-494, 1441-1462): (vp, vs, rho[, delta, eps]) -> (lambda, mu, rho[, c13, eps]);
// the acoustic nl_we_op_a::convert_model (we_op.cpp:1863-1881): (vp, rho) -> (K = rho vp^2, 1/rho)
void convert_model(std::vector<double>& m, int np, long long n)
{
    if (np == 2) {
        for (long long i = 0; i < n; i++) {
            m[i] = m[n + i] * m[i] * m[i];
            m[n + i] = 1.0 / m[n + i];
        }
        return;
    }
    for (long long i = 0; i < n; i++) {
        const double vp = m[i], vs = m[n + i], rho = m[2 * n + i];
        const double la = rho * (vp * vp - 2 * vs * vs), mu = rho * vs * vs;
        m[i] = la;
        m[n + i] = mu;
        if (np == 5) m[3 * n + i] = std::sqrt(2 * (la + 2 * mu) * (la + mu) * m[3 * n + i] + (la + mu) * (la + mu)) - mu;
    }
}

// sinc_resampler (operator.cpp:808-856, alpha = 0.5) and linear_resampler (760-806):
// model axis (ntm, dtm) -> data axis (ntd, dtd), both with origin 0.
void resample(bool adjoint, bool linear, int hl, int ntr, int ntm, double dtm, int ntd, double dtd, double* mod,
              double* dat)
{
    if (adjoint) std::fill(mod, mod + (size_t)ntr * ntm, 0.0);
    for (int x = 0; x < ntr; x++) {
        double* pm = mod + (size_t)x * ntm;
        double* pd = dat + (size_t)x * ntd;
        for (int itd = 0; itd < ntd; itd++) {
            const double t = itd * dtd;
            if (linear) {
                const int i = std::min(ntm - 1, (int)std::floor(t / dtm));
                const int j = std::min(ntm - 1, (int)std::ceil(t / dtm));
                const double wj = (t - i * dtm) / dtm, wi = 1 - wj;
                if (!adjoint) pd[itd] = wi * pm[i] + wj * pm[j];
                else { pm[i] += wi * pd[itd]; pm[j] += wj * pd[itd]; }
                continue;
            }
            const int itm0 = (int)std::floor(t / dtm);
            const int lo = std::max(0, itm0 - hl + 1), hi = std::min(ntm - 1, itm0 + hl);
            double v = 0;
            for (int itm = lo; itm <= hi; itm++) {
                const double arg = kPi * (t - itm * dtm) / dtm;
                const double s = arg == 0.0 ? 1.0 : std::sin(arg) / arg;
                const double w = s * (0.5 + 0.5 * std::cos(kPi * (itm - t / dtm) / hl));
                if (!adjoint) v += pm[itm] * w;
                else pm[itm] += pd[itd] * w;
            }
            if (!adjoint) pd[itd] = v;
        }
    }
}

// Per-trace linear maps in time as sparse rows (row i: (column, weight) pairs), composed
// on the host in double precision into one TraceMatrix that the device applies.
using Rows = std::vector<std::vector<std::pair<int, double>>>;

// the resampler as a matrix [ntd][ntm]: dat = S mod (resample above, forward direction)
Rows resample_rows(bool linear, int hl, int ntm, double dtm, int ntd, double dtd)
{
    Rows r(ntd);
    for (int itd = 0; itd < ntd; itd++) {
        const double t = itd * dtd;
        if (linear) {
            const int i = std::min(ntm - 1, (int)std::floor(t / dtm));
            const int j = std::min(ntm - 1, (int)std::ceil(t / dtm));
            const double wj = (t - i * dtm) / dtm, wi = 1 - wj;
            if (i == j) r[itd] = {{i, wi + wj}};
            else r[itd] = {{i, wi}, {j, wj}};
            continue;
        }
        const int itm0 = (int)std::floor(t / dtm);
        const int lo = std::max(0, itm0 - hl + 1), hi = std::min(ntm - 1, itm0 + hl);
        for (int itm = lo; itm <= hi; itm++) {
            const double arg = kPi * (t - itm * dtm) / dtm;
            const double sc = arg == 0.0 ? 1.0 : std::sin(arg) / arg;
            r[itd].push_back({itm, sc * (0.5 + 0.5 * std::cos(kPi * (itm - t / dtm) / hl))});
        }
    }
    return r;
}

// Dt (misc.cpp:71-95) as a matrix [nt][nt]: the centred derivative with one-sided ends,
// or the reference's reverse-time adjoint form
Rows dt_rows(bool adj, int nt, double dt)
{
    Rows r(nt);
    const double a = 1 / dt, b = 1 / (2 * dt);
    r[0] = adj ? Rows::value_type{{0, a}, {1, a}} : Rows::value_type{{0, -a}, {1, a}};
    for (int t = 1; t < nt - 1; t++) r[t] = {{t - 1, -b}, {t + 1, b}};
    r[nt - 1] = adj ? Rows::value_type{{nt - 2, -a}, {nt - 1, -a}} : Rows::value_type{{nt - 2, -a}, {nt - 1, a}};
    return r;
}

Rows transpose(const Rows& a, int ncol)
{
    Rows t(ncol);
    for (int i = 0; i < (int)a.size(); i++)
        for (const auto& e : a[i]) t[e.first].push_back({i, e.second});
    return t;
}

// a b, with b of bcols columns
Rows multiply(const Rows& a, const Rows& b, int bcols)
{
    Rows c(a.size());
    std::vector<double> acc(bcols, 0.0);
    std::vector<char> used(bcols, 0);
    std::vector<int> cols;
    for (size_t i = 0; i < a.size(); i++) {
        cols.clear();
        for (const auto& e : a[i])
            for (const auto& f : b[e.first]) {
                if (!used[f.first]) {
                    used[f.first] = 1;
                    cols.push_back(f.first);
                }
                acc[f.first] += e.second * f.second;
            }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) {
            c[i].push_back({j, acc[j]});
            acc[j] = 0.0;
            used[j] = 0;
        }
    }
    return c;
}

fwi2d::TraceMatrix to_matrix(const Rows& a, int ncol)
{
    fwi2d::TraceMatrix m;
    m.nrow = (int)a.size();
    m.ncol = ncol;
    m.ptr.assign(1, 0);
    for (const auto& row : a) {
        for (const auto& e : row) {
            m.col.push_back(e.first);
            m.w.push_back(e.second);
        }
        m.ptr.push_back((int)m.col.size());
    }
    return m;
}

double h_weight(int n, int i)
{
    static const double w[4] = {17.0 / 48, 59.0 / 48, 43.0 / 48, 49.0 / 48};
    if (i < 4) return w[i];
    if (i >= n - 4) return w[n - 1 - i];
    return 1.0;
}

// f(x, ws) for x in [0, n) on up to 16 host threads, in chunks of 64 taken in turn; ws is
// a workspace of type W that each thread constructs once and reuses.  Every x is computed
// the same way whatever thread takes it.  An exception thrown by f (or by a thread's
// start) stops the hand-out of further chunks and is rethrown here once every started
// thread has joined, so it reaches the caller's error handling instead of terminating
// the process.
template <class W, class F>
void parallel_traces(size_t n, F&& f)
{
    const unsigned nth = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 256 || nth == 1) {
        W ws;
        for (size_t x = 0; x < n; x++) f(x, ws);
        return;
    }
    std::atomic<size_t> next{0};
    std::mutex mu;
    std::exception_ptr err;
    auto work = [&] {
        try {
            W ws;
            for (size_t b; (b = next.fetch_add(64)) < n;)
                for (size_t x = b; x < std::min(n, b + 64); x++) f(x, ws);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!err) err = std::current_exception();
            next.store(n);
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(nth);
        for (unsigned t = 0; t < nth; t++) pool.emplace_back(work);
    } catch (...) {
        next.store(n);
        for (auto& t : pool) t.join();
        throw;
    }
    for (auto& t : pool) t.join();
    if (err) std::rethrow_exception(err);
}

// conv1dnd (operator.hpp:1035-1057; operator.cpp:2092-2137) of one trace of n samples
// with the filter f: y[t] = sum_j f[j] x[t - j + sh] over the j that keep t - j + sh in
// the trace (sh = floor((nf + 1)/2) - 1 centred, else 0); the adjoint scatters
// y[t] f[j] to x[t - j + sh].  x and y must not overlap.
void conv_trace(const std::vector<double>& f, bool centered, bool adjoint, const double* x, double* y, int n)
{
    const int nf = (int)f.size();
    const int sh = centered ? (nf + 1) / 2 - 1 : 0;
    std::fill(y, y + n, 0.0);
    for (int j = 0; j < nf; j++) {
        const int k = sh - j;  // forward: y[t] takes x[t + k]
        const int lo = std::max(0, -k), hi = std::min(n, n - k);
        const double fj = f[j];
        if (!adjoint)
            for (int t = lo; t < hi; t++) y[t] += fj * x[t + k];
        else
            for (int t = lo; t < hi; t++) y[t + k] += fj * x[t];
    }
}

// In-place iterative radix-2 complex FFT of n = 2^k points (re, im), forward (sign -1)
// or inverse without the 1/n (sign +1); tw holds exp(-2 pi i j / n), j < n/2.
void fft_pow2(std::vector<double>& re, std::vector<double>& im, const std::vector<double>& twr,
              const std::vector<double>& twi, bool inverse)
{
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; i++) {  // bit reversal
        size_t b = n >> 1;
        for (; j & b; b >>= 1) j ^= b;
        j |= b;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; k++) {
                const double wr = twr[k * step], wi = inverse ? -twi[k * step] : twi[k * step];
                const size_t a = i + k, b = a + half;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

// conv1dnd of traces of n samples with one filter, as conv_trace computes it: directly for
// short filters, through zero-padded FFTs of 2^k >= n + nf - 1 points for long ones (the
// circular convolution is then the linear one; the adjoint is the circular correlation
// with the filter, read at lag k - sh).  Both are exact up to double rounding.
struct TraceFilter {
    std::vector<double> f;
    bool centered = true;
    int n = 0;
    size_t N = 0;
    std::vector<double> twr, twi, fr, fi;  // twiddles; the filter's spectrum
    struct Work {
        std::vector<double> re, im;
    };

    void init(const std::vector<double>& filt, bool cen, int nt)
    {
        f = filt;
        centered = cen;
        n = nt;
        N = 0;
        if (f.size() < 48) return;  // direct sums are cheaper
        N = 1;
        while (N < (size_t)n + f.size() - 1) N <<= 1;
        twr.resize(N / 2);
        twi.resize(N / 2);
        for (size_t j = 0; j < N / 2; j++) {
            twr[j] = std::cos(2 * kPi * (double)j / (double)N);
            twi[j] = -std::sin(2 * kPi * (double)j / (double)N);
        }
        fr.assign(N, 0.0);
        fi.assign(N, 0.0);
        std::copy(f.begin(), f.end(), fr.begin());
        fft_pow2(fr, fi, twr, twi, false);
    }

    bool empty() const { return f.empty(); }

    // y = C x (adjoint: C^T x); x and y must not overlap
    void apply(bool adjoint, const double* x, double* y, Work& w) const { apply2(adjoint, x, y, nullptr, nullptr, w); }

    // y0 = C x0 and, when x1 is given, y1 = C x1, with one complex FFT pair for both on the
    // FFT path: x0 + i x1 is transformed, and since the filter is real the real and
    // imaginary parts of the filtered signal are the two filtered traces.  No output may
    // overlap an input.
    void apply2(bool adjoint, const double* x0, double* y0, const double* x1, double* y1, Work& w) const
    {
        if (!N) {
            conv_trace(f, centered, adjoint, x0, y0, n);
            if (x1) conv_trace(f, centered, adjoint, x1, y1, n);
            return;
        }
        const int sh = centered ? ((int)f.size() + 1) / 2 - 1 : 0;
        w.re.assign(N, 0.0);
        w.im.assign(N, 0.0);
        std::copy(x0, x0 + n, w.re.begin());
        if (x1) std::copy(x1, x1 + n, w.im.begin());
        fft_pow2(w.re, w.im, twr, twi, false);
        for (size_t k = 0; k < N; k++) {
            const double a = w.re[k], b = w.im[k], c = fr[k], d = adjoint ? -fi[k] : fi[k];
            w.re[k] = a * c - b * d;
            w.im[k] = a * d + b * c;
        }
        fft_pow2(w.re, w.im, twr, twi, true);
        const double s = 1.0 / (double)N;
        for (int t = 0; t < n; t++) {
            const long long m = adjoint ? ((long long)t - sh + (long long)N) % (long long)N : (long long)t + sh;
            y0[t] = w.re[(size_t)m] * s;
            if (x1) y1[t] = w.im[(size_t)m] * s;
        }
    }
};

// per-thread scratch of the objective's trace loop (a pair of traces at a time)
struct TraceWork {
    std::vector<double> v[2], s[2], c[2];
    TraceFilter::Work fw;
};

// Real-input discrete Fourier transform of x (n values) onto its n/2 + 1 non-negative
// frequencies, and the inverse from those frequencies (with 1/n, the imaginary part of
// frequency 0 ignored, as FFTW's r2c / c2r pair that fxTransform uses, operator.cpp:
// 858-977).  Plain O(n^2) sums in double precision: they run once per filter.
void dft_r2c(const std::vector<double>& x, std::vector<double>& re, std::vector<double>& im)
{
    const int n = (int)x.size(), nf = n / 2 + 1;
    re.assign(nf, 0.0);
    im.assign(nf, 0.0);
    for (int k = 0; k < nf; k++) {
        double a = 0, b = 0;
        for (int t = 0; t < n; t++) {
            const double ph = 2 * kPi * (double)(((long long)k * t) % n) / n;
            a += x[t] * std::cos(ph);
            b -= x[t] * std::sin(ph);
        }
        re[k] = a;
        im[k] = b;
    }
}

std::vector<double> dft_c2r(const std::vector<double>& re, const std::vector<double>& im, int n)
{
    std::vector<double> x(n);
    const int nf = n / 2 + 1;
    for (int t = 0; t < n; t++) {
        double s = re[0];
        for (int k = 1; k < nf; k++) {
            // the conjugate-symmetric upper half counts twice, except a Nyquist frequency
            const double m = (2 * k == n) ? 1.0 : 2.0;
            const double ph = 2 * kPi * (double)(((long long)k * t) % n) / n;
            s += m * (re[k] * std::cos(ph) - im[k] * std::sin(ph));
        }
        x[t] = s / n;
    }
    return x;
}

// The filter as nlls_fwi_eco's constructor prepares it (optimization.hpp:526-534), padded
// to odd length for the zero and minimum phases: zero_phase (operator.cpp:2197-2250),
// the amplitude spectrum with the linear phase exp(-i w (n-1)/2); minimum_phase
// (operator.cpp:2252-2320), exp(FT(causal fold of the real cepstrum of log(|FT f| + eps)))
// with eps = 1e-7.
std::vector<double> prepare_filter(const float* f, int nf, int phase)
{
    std::vector<double> x(f, f + nf);
    if (phase == FWI2D_FILTER_DEFAULT) return x;
    if (nf % 2 == 0) x.push_back(0.0);
    const int n = (int)x.size(), nh = n / 2 + 1;
    std::vector<double> re, im;
    dft_r2c(x, re, im);
    if (phase == FWI2D_FILTER_ZERO) {
        for (int k = 0; k < nh; k++) {
            const double r = std::hypot(re[k], im[k]);
            const double ph = -k * (2 * kPi / n) * (n - 1) / 2;
            re[k] = r * std::cos(ph);
            im[k] = r * std::sin(ph);
        }
        return dft_c2r(re, im, n);
    }
    for (int k = 0; k < nh; k++) {
        re[k] = std::log(std::hypot(re[k], im[k]) + 1e-7);
        im[k] = 0.0;
    }
    std::vector<double> c = dft_c2r(re, im, n);
    for (int k = 1; k < nh; k++) c[k] *= 2;
    for (int k = nh; k < n; k++) c[k] = 0.0;
    dft_r2c(c, re, im);
    for (int k = 0; k < nh; k++) {
        const double a = std::exp(re[k]);
        re[k] = a * std::cos(im[k]);
        im[k] = a * std::sin(im[k]);
    }
    return dft_c2r(re, im, n);
}

}  // namespace

struct fwi2d_op {
    fwi2d_op_desc d{};
    std::vector<double> sxz, rxz, rdip;
    std::vector<int> nrec;
    int nr = 0;
    std::vector<double> allsrc;  // [nc][ns][nt_data], nc = 2 (forces), 3 (moment tensors), 1 (acoustic)
    fwi2d_engine* eng = nullptr;
    fwi2d_acoustic* aeng = nullptr;  // the acoustic operator's engine (nparam = 2)
    fwi2d_grid grid{};
    TimeAxis axis{0, 0, 0};
    std::vector<float> ckpt_model;  // converted model of the last checkpointed forward
    std::vector<float> src_prop;    // source traces at the propagation step
    fwi2d::TraceMatrix tm_fwd;      // receiver traces -> data [nt_data][nt]
    fwi2d::TraceMatrix tm_adj;      // data -> adjoint sources [nt][nt_data]
    bool shots_set = false;

    int np() const { return d.nparam; }
    bool acoustic() const { return d.nparam == 2; }
    int nc() const { return acoustic() ? 1 : (d.mt ? 3 : 2); }  // source components
    int ncd() const { return acoustic() ? 1 : 2; }               // data components
    long long nxz() const { return (long long)d.nx * d.nz; }

    // model (physical) -> analysed + converted float model; sets the time axis
    std::vector<float> prepare(const float* pmod)
    {
        const long long n = nxz();
        std::vector<double> m(pmod, pmod + (size_t)np() * n);
        for (double v : m)
            if (!std::isfinite(v)) throw std::runtime_error("model contains a non-finite value");
        TimeAxis t = analyze_model(d, m);
        convert_model(m, np(), n);
        ensure_engine(t);
        return std::vector<float>(m.begin(), m.end());
    }

    void ensure_engine(const TimeAxis& t)
    {
        if ((eng || aeng) && t.nt == axis.nt && t.dt == axis.dt && t.sub == axis.sub) return;
        if (eng) fwi2d_engine_destroy(eng);
        if (aeng) fwi2d_acoustic_destroy(aeng);
        eng = nullptr;
        aeng = nullptr;
        axis = t;
        fwi2d_grid g{};
        g.nx = d.nx;
        g.nz = d.nz;
        g.dx = d.dx;
        g.dz = d.dz;
        g.nt = t.nt;
        g.dt = t.dt;
        g.sub = t.sub;
        g.bc_top = d.bc_top;
        g.bc_bottom = d.bc_bottom;
        g.bc_left = d.bc_left;
        g.bc_right = d.bc_right;
        g.taper_top = d.taper_top;
        g.taper_bottom = d.taper_bottom;
        g.taper_left = d.taper_left;
        g.taper_right = d.taper_right;
        g.taper_strength = d.taper_strength;
        g.vti = d.nparam == 5;
        g.gdamp_tmin = d.gdamp_tmin;
        g.gdamp_tmax = d.gdamp_tmax;
        g.gdamp_power = d.gdamp_power;
        grid = g;
        if (acoustic()) {
            check(fwi2d_acoustic_create(&g, d.free_surface_stiffness, d.device, &aeng));
            check(fwi2d_acoustic_set_shots(aeng, d.ns, sxz.data(), nrec.data(), rxz.data(), d.ox, d.oz));
        } else {
            check(fwi2d_engine_create(&g, d.device, &eng));
            check(fwi2d_engine_set_acquisition(eng, d.ns, sxz.data(), nrec.data(), rxz.data(), d.ox, d.oz, d.mt,
                                               d.gl > 0 ? d.gl : 0.0, d.gl > 0 ? rdip.data() : nullptr));
        }
        ckpt_model.clear();
        // sources at the propagation step (apply_forward, we_op.cpp:953-958)
        std::vector<double> sp((size_t)nc() * d.ns * t.nt);
        resample(false, d.resampling == 1, d.sinc_half_length, nc() * d.ns, d.nt_data, d.dt_data, t.nt, t.dt,
                 allsrc.data(), sp.data());
        src_prop.assign(sp.begin(), sp.end());
        build_trace_maps();
    }

    // The trace maps, applied on the device to every trace: apply_forward's (we_op.cpp:
    // 1000-1030) alpha Dt (particle velocity) then the resampler's adjoint onto the data
    // axis, and the adjoint sources of apply_jacobianT (we_op.cpp:1112-1127): resampling
    // onto the propagation axis, Ht^-1, time reversal, -alpha, then Dt's adjoint form.
    void build_trace_maps()
    {
        const int nt = axis.nt, ntd = d.nt_data;
        const double dt = axis.dt, alpha = dt / d.dt_data;
        const Rows S = resample_rows(d.resampling == 1, d.sinc_half_length, ntd, d.dt_data, nt, dt);  // [nt][ntd]
        Rows F = transpose(S, ntd);
        if (d.seismotype == 1) F = multiply(F, dt_rows(false, nt, dt), nt);
        for (auto& row : F)
            for (auto& e : row) e.second *= alpha;
        tm_fwd = to_matrix(F, nt);
        Rows A(nt);
        for (int t = 0; t < nt; t++) {
            const int k = nt - 1 - t;  // sample k of the resampled trace lands at t
            const double h = (k == 0 || k == nt - 1) ? 0.5 * dt : dt;
            A[t] = S[k];
            for (auto& e : A[t]) e.second *= -alpha / h;
        }
        if (d.seismotype == 1) A = multiply(dt_rows(true, nt, dt), A, ntd);
        tm_adj = to_matrix(A, ntd);
    }

    void set_model(const std::vector<float>& m)
    {
        if (acoustic()) check(fwi2d_acoustic_set_model(aeng, m.data()));
        else check(fwi2d_engine_set_model(eng, m.data()));
    }

    void run_forward(const std::vector<float>& m, std::vector<float>& rcv)
    {
        set_model(m);
        rcv.assign((size_t)ncd() * nr * axis.nt, 0.f);
        const int keep = axis.sub > 0;
        if (acoustic()) check(fwi2d_acoustic_forward(aeng, src_prop.data(), rcv.data(), keep, 0));
        else check(fwi2d_engine_forward(eng, src_prop.data(), rcv.data(), keep, 0));
        if (keep) ckpt_model = m;
        else ckpt_model.clear();
    }

    void apply_forward(bool add, const float* pmod, float* pdat)
    {
        std::vector<float> m = prepare(pmod);
        if (nr == 0) {  // no traces: the run only leaves checkpoints
            std::vector<float> rcv;
            run_forward(m, rcv);
            return;
        }
        set_model(m);
        ckpt_model.clear();
        const size_t n = (size_t)ncd() * nr * d.nt_data;
        std::vector<float> out(n);
        const bool keep = axis.sub > 0;
        if (acoustic()) fwi2d::acoustic_forward_traces(aeng, src_prop.data(), keep, tm_fwd, out.data());
        else fwi2d::engine_forward_traces(eng, src_prop.data(), keep, tm_fwd, out.data());
        if (keep) ckpt_model = m;
        for (size_t i = 0; i < n; i++) pdat[i] = add ? pdat[i] + out[i] : out[i];
    }

    void apply_jacobianT(bool add, float* pmod, const float* pmod0, const float* pdat)
    {
        std::vector<float> m = prepare(pmod0);
        if (axis.sub <= 0) throw std::runtime_error("apply_jacobianT needs sub != 0");
        const long long n = nxz();
        const int ng = acoustic() ? 2 : 3;  // raw gradient fields
        std::vector<float> graw((size_t)(acoustic() ? 2 : np()) * n, 0.f);
        if (nr > 0) {  // no traces: no adjoint sources, a zero gradient
            if (ckpt_model != m) {  // no checkpoints of this model: run the forward first
                std::vector<float> rcv;
                run_forward(m, rcv);
            }
            // (DAS channels: the device maps strain to gauge-end differences first,
            // dipole_to_strain's adjoint; it commutes with the per-trace time map)
            if (acoustic()) fwi2d::acoustic_adjoint_traces(aeng, pdat, tm_adj, graw.data());
            else fwi2d::engine_adjoint_traces(eng, pdat, tm_adj, graw.data());
        }
        // H quadrature (we_op.cpp:1192-1199) and chain rule (we_op.cpp:1208-1230)
        std::vector<double> gd(graw.begin(), graw.end());
        for (int c = 0; c < ng; c++)
            for (int ix = 0; ix < d.nx; ix++)
                for (int iz = 0; iz < d.nz; iz++)
                    gd[c * n + (long long)ix * d.nz + iz] *= h_weight(d.nz, iz) * d.dz * h_weight(d.nx, ix) * d.dx;
        if (acoustic()) {
            // (K, 1/rho) -> (vp, rho): g_rho = (1/rho) K g_K - (1/rho)^2 g_1/rho, g_vp = 2 sqrt(K rho) g_K
            for (long long i = 0; i < n; i++) {
                const double K = m[i], ri = m[n + i];
                const double gk = gd[i], gri = gd[n + i];
                gd[n + i] = ri * K * gk - ri * ri * gri;
                gd[i] = 2 * std::sqrt(K / ri) * gk;
            }
        } else {
            for (long long i = 0; i < n; i++) {
                const double la = m[i], mu = m[n + i], rho = m[2 * n + i];
                const double gl = gd[i], gm = gd[n + i], gr = gd[2 * n + i];
                gd[2 * n + i] = (la / rho) * gl + (mu / rho) * gm + gr;
                gd[n + i] = 2 * std::sqrt(mu * rho) * (gm - 2 * gl);
                gd[i] = 2 * std::sqrt(rho * (la + 2 * mu)) * gl;
            }
            for (long long i = 3 * n; i < np() * n; i++) gd[i] = 0.0;
        }
        for (long long i = 0; i < np() * n; i++) pmod[i] = (float)((add ? (double)pmod[i] : 0.0) + gd[i]);
    }

    ~fwi2d_op()
    {
        if (eng) fwi2d_engine_destroy(eng);
        if (aeng) fwi2d_acoustic_destroy(aeng);
    }
};

// Least-squares FWI objective on the operator: nlls_fwi (optimization.hpp:295-455) and
// nlls_fwi_eco (optimization.hpp:459-951) with residual weights, time integration, a 1-D
// filter, trace normalisation, a gradient mask and nlls_fwi_eco's data normalisation (no
// envelope, interferometry, double difference or source rescaling).  Both shot loops
// become two batched phases: one apply_forward over every shot, the residuals, one
// apply_jacobianT(add=false) over every shot.  That is the sum of nlls_fwi_eco's per-shot
// apply_jacobianT(add=true) calls (optimization.hpp:890), without the CPU build's
// re-transform of the running sum (we_op.cpp:1080-1084, 1192-1236).  The trace
// processing and the reductions run in double precision (the reference's in data_t).
struct fwi2d_nlls {
    fwi2d_op* op = nullptr;
    bool integrate = false, trace_norm = false;
    std::vector<double> d;     // processed observed data: integrated, weighted, normalised
    std::vector<float> w;      // residual weights, empty: none
    std::vector<float> gmask;  // gradient mask, empty: none
    TraceFilter filt;          // prepared filter, empty: none
    std::vector<float> r;      // residual of the last evaluation
    double dnorm = 1.0;        // nlls_fwi_eco's data norm (optimization.hpp:569-574), else 1
    size_t ndata = 0, nmodel = 0;

    int nt() const { return op->d.nt_data; }
    size_t ntr() const { return ndata / nt(); }

    // integral (operator.hpp:650-689) of one trace: trapezoid running sum, or its adjoint
    void integral_trace(double* p, bool adjoint) const
    {
        const int n = nt();
        const double dt = op->d.dt_data;
        if (!adjoint) {
            p[0] = 0.5 * p[0] * dt;
            for (int t = 1; t < n - 1; t++) p[t] = p[t - 1] + p[t] * dt;
            p[n - 1] = p[n - 2] + 0.5 * p[n - 1] * dt;
        } else {
            p[n - 1] = p[n - 1] * dt;
            for (int t = n - 2; t >= 0; t--) p[t] = p[t + 1] + p[t] * dt;
            p[n - 1] *= 0.5;
            p[0] *= 0.5;
        }
    }

    // ttnormalize (misc.cpp:97-110) of one trace: divided by its L2 norm (1 below 1e-16),
    // which is returned
    double normalize_trace(double* p) const
    {
        const int n = nt();
        double s = 0;
        for (int t = 0; t < n; t++) s += p[t] * p[t];
        if (s <= 1e-16) s = 1;
        s = std::sqrt(s);
        for (int t = 0; t < n; t++) p[t] /= s;
        return s;
    }

    // the data processing of the constructors (optimization.hpp:512-542): integrate,
    // weight, filter, normalise
    void process(std::vector<double>& u) const
    {
        const int n = nt();
        const size_t T = ntr();
        parallel_traces<TraceWork>((T + 1) / 2, [&](size_t pr, TraceWork& ws) {
            const int m = 2 * pr + 1 < T ? 2 : 1;
            double* p[2] = {u.data() + 2 * pr * n, m == 2 ? u.data() + (2 * pr + 1) * n : nullptr};
            for (int q = 0; q < m; q++) {
                if (integrate) integral_trace(p[q], false);
                if (!w.empty())
                    for (int t = 0; t < n; t++) p[q][t] *= w[(2 * pr + q) * n + t];
            }
            if (!filt.empty()) {
                ws.c[0].resize(n);
                ws.c[1].resize(n);
                filt.apply2(false, p[0], ws.c[0].data(), p[1], ws.c[1].data(), ws.fw);
                for (int q = 0; q < m; q++) std::copy(ws.c[q].begin(), ws.c[q].end(), p[q]);
            }
            if (trace_norm)
                for (int q = 0; q < m; q++) normalize_trace(p[q]);
        });
    }

    void evaluate(const float* m, double* func, float* g)
    {
        const int n = nt();
        const double dt = op->d.dt_data;
        const size_t T = ntr();
        std::vector<float> u(ndata);
        op->apply_forward(false, m, u.data());
        r.resize(ndata);
        std::vector<double> fpart(T);
        std::vector<float> af(g ? ndata : 0);
        // trace by trace (each independent, so the result does not depend on the threads):
        // res (optimization.hpp:346-374, 660-776): r = P(F(m)) - P(d), P the processing;
        // the trace's share of getFunc (optimization.hpp:375-385, 929-942):
        // |r|^2 - 1/2 (r[0]^2 + r[n-1]^2); and the adjoint source of grad
        // (optimization.hpp:390-455, 780-905): I^T W C^T N'^T Ht r, with
        // N' = I/|s| - s s^T/|s|^3 the Jacobian of the trace normalisation at s
        // Traces go in pairs (2p, 2p + 1), whose filters share one complex FFT pair.
        parallel_traces<TraceWork>((T + 1) / 2, [&](size_t pr, TraceWork& ws) {
            const int m = 2 * pr + 1 < T ? 2 : 1;
            double nrm[2] = {1.0, 1.0};
            auto filter_pair = [&](bool adjoint) {
                ws.c[0].resize(n);
                ws.c[1].resize(n);
                filt.apply2(adjoint, ws.v[0].data(), ws.c[0].data(), m == 2 ? ws.v[1].data() : nullptr,
                            ws.c[1].data(), ws.fw);
                for (int q = 0; q < m; q++) ws.v[q].swap(ws.c[q]);
            };
            for (int q = 0; q < m; q++) {
                const size_t x = 2 * pr + q;
                std::vector<double>& v = ws.v[q];
                v.assign(u.begin() + x * n, u.begin() + (x + 1) * n);
                if (integrate) integral_trace(v.data(), false);
                if (!w.empty())
                    for (int t = 0; t < n; t++) v[t] *= w[x * n + t];
            }
            if (!filt.empty()) filter_pair(false);
            for (int q = 0; q < m; q++) {
                const size_t x = 2 * pr + q;
                std::vector<double>& v = ws.v[q];
                std::vector<double>& s = ws.s[q];
                if (trace_norm) {
                    s = v;
                    nrm[q] = normalize_trace(v.data());
                }
                float* rx = r.data() + x * n;
                const double* dx = d.data() + x * n;
                double f = 0;
                for (int t = 0; t < n; t++) {
                    rx[t] = (float)(v[t] - dx[t]);
                    f += (double)rx[t] * rx[t];
                }
                fpart[x] = f - 0.5 * ((double)rx[0] * rx[0] + (double)rx[n - 1] * rx[n - 1]);
                if (!g) continue;
                for (int t = 0; t < n; t++) v[t] = (double)rx[t] * ((t == 0 || t == n - 1) ? 0.5 * dt : dt);
                if (trace_norm) {
                    double vs = 0;
                    for (int t = 0; t < n; t++) vs += v[t] * s[t];
                    for (int t = 0; t < n; t++) v[t] = v[t] / nrm[q] - vs / (nrm[q] * nrm[q] * nrm[q]) * s[t];
                }
            }
            if (!g) return;
            if (!filt.empty()) filter_pair(true);
            for (int q = 0; q < m; q++) {
                const size_t x = 2 * pr + q;
                std::vector<double>& v = ws.v[q];
                if (!w.empty())
                    for (int t = 0; t < n; t++) v[t] *= w[x * n + t];
                if (integrate) integral_trace(v.data(), true);
                for (int t = 0; t < n; t++) af[x * n + t] = (float)v[t];
            }
        });
        if (func) {
            double f = 0;
            for (double p : fpart) f += p;
            *func = 0.5 * dt * f / dnorm;
        }
        if (!g) return;
        std::vector<float> gr(nmodel);
        op->apply_jacobianT(false, gr.data(), m, af.data());
        for (size_t i = 0; i < nmodel; i++)
            g[i] = (float)((gmask.empty() ? 1.0 : (double)gmask[i]) * (gr[i] / dnorm));
    }
};

template <class F>
static int op_guard(F&& f)
{
    try {
        f();
        fwi2d::g_last_error.clear();
        return 0;
    } catch (const std::exception& ex) {
        fwi2d::g_last_error = ex.what();
        return 1;
    }
}

extern "C" {

int fwi2d_op_create(const fwi2d_op_desc* desc, const float* allsrc, fwi2d_op** out)
{
    return op_guard([&] {
        if (!desc || !allsrc || !out) throw std::runtime_error("null argument");
        const fwi2d_op_desc& d = *desc;
        if (d.nparam != 2 && d.nparam != 3 && d.nparam != 5)
            throw std::runtime_error("The model must contain two (Vp, rho), three (Vp, Vs, rho) or five (Vp, Vs, rho, "
                                     "delta, eps) parameters");
        if (d.ns < 1) throw std::runtime_error("need at least one shot");
        if (d.nt_data < 2 || !(d.dt_data > 0)) throw std::runtime_error("invalid data time axis");
        fwi2d_op* op = new fwi2d_op;
        op->d = d;
        if (d.nparam == 2) {  // the acoustic operator: point sources and receivers (we_op.cpp:13-20)
            op->d.mt = 0;
            op->d.gl = 0;
        }
        op->sxz.assign(d.sxz, d.sxz + 2 * d.ns);
        op->nrec.assign(d.nrec, d.nrec + d.ns);
        op->nr = 0;
        for (int r : op->nrec) op->nr += r;
        op->rxz.assign(d.rxz, d.rxz + 2 * (size_t)op->nr);
        if (op->d.gl > 0) {
            if (!d.rdip) {
                delete op;
                throw std::runtime_error("DAS receivers (gl > 0) need their dips (rdip)");
            }
            op->rdip.assign(d.rdip, d.rdip + op->nr);
        }
        op->d.sxz = nullptr;
        op->d.nrec = nullptr;
        op->d.rxz = nullptr;
        op->d.rdip = nullptr;
        if (op->d.sinc_half_length <= 0) op->d.sinc_half_length = 11;
        // shot and receiver positions inside the model (analyzeGeometry, we_op.cpp:98-127)
        const double prec = std::min(d.dx, d.dz) * 1e-3;
        const double xmax = d.ox + (d.nx - 1) * d.dx, zmax = d.oz + (d.nz - 1) * d.dz;
        auto inside = [&](double x, double z) {
            return x - d.ox >= -prec && -prec <= xmax - x && z - d.oz >= -prec && -prec <= zmax - z;
        };
        for (int s = 0; s < d.ns; s++)
            if (!inside(op->sxz[2 * s], op->sxz[2 * s + 1])) {
                delete op;
                throw std::runtime_error("Shot falls outside subsurface model boundaries");
            }
        for (int r = 0; r < op->nr; r++) {
            const double x = op->rxz[2 * r], z = op->rxz[2 * r + 1];
            if (op->d.gl <= 0 && !inside(x, z)) {
                delete op;
                throw std::runtime_error("Receiver " + std::to_string(r) + " falls outside subsurface model boundaries");
            }
            if (op->d.gl > 0) {  // both gauge ends inside the model (we_op.cpp:134-148)
                const double hx = d.gl / 2 * std::cos(op->rdip[r]), hz = d.gl / 2 * std::sin(op->rdip[r]);
                const bool out = x - hx < d.ox || x - hx > xmax || x + hx < d.ox || x + hx > xmax || z - hz < d.oz ||
                                 z - hz > zmax || z + hz < d.oz || z + hz > zmax;
                if (out) {
                    delete op;
                    throw std::runtime_error("DAS channel " + std::to_string(r) + " must be away from the subsurface "
                                             "model boundaries by half of the gauge length along the fiber");
                }
            }
        }
        op->allsrc.assign(allsrc, allsrc + (size_t)op->nc() * d.ns * d.nt_data);
        *out = op;
    });
}

int fwi2d_op_apply_forward(fwi2d_op* op, int add, const float* pmod, float* pdat)
{
    return op_guard([&] {
        if (!op || !pmod || !pdat) throw std::runtime_error("null argument");
        op->apply_forward(add != 0, pmod, pdat);
    });
}

int fwi2d_op_apply_jacobianT(fwi2d_op* op, int add, float* pmod, const float* pmod0, const float* pdat)
{
    return op_guard([&] {
        if (!op || !pmod || !pmod0 || !pdat) throw std::runtime_error("null argument");
        op->apply_jacobianT(add != 0, pmod, pmod0, pdat);
    });
}

fwi2d_engine* fwi2d_op_engine(fwi2d_op* op) { return op ? op->eng : nullptr; }

int fwi2d_op_time_axis(fwi2d_op* op, int* nt, double* dt, int* sub)
{
    return op_guard([&] {
        if (!op) throw std::runtime_error("null argument");
        if (nt) *nt = op->axis.nt;
        if (dt) *dt = op->axis.dt;
        if (sub) *sub = op->axis.sub;
    });
}

void fwi2d_op_destroy(fwi2d_op* op) { delete op; }

int fwi2d_nlls_create(fwi2d_op* op, const float* d, const float* w, const float* gmask, int flags,
                      fwi2d_nlls** out)
{
    return fwi2d_nlls_create_filtered(op, d, w, gmask, flags, nullptr, 0, FWI2D_FILTER_DEFAULT, out);
}

int fwi2d_nlls_create_filtered(fwi2d_op* op, const float* d, const float* w, const float* gmask, int flags,
                               const float* filter, int nf, int filter_phase, fwi2d_nlls** out)
{
    return op_guard([&] {
        if (!op || !d || !out) throw std::runtime_error("null argument");
        if (flags & ~(FWI2D_NLLS_NORMALIZE_OBJ | FWI2D_NLLS_INTEGRATE | FWI2D_NLLS_TRACE_NORMALIZE))
            throw std::runtime_error("unknown objective flags");
        if (filter && nf < 1) throw std::runtime_error("the filter needs at least one sample");
        if (filter_phase < FWI2D_FILTER_DEFAULT || filter_phase > FWI2D_FILTER_MINIMUM)
            throw std::runtime_error("unknown filter phase");
        const size_t ndata = (size_t)op->ncd() * op->nr * op->d.nt_data, nmodel = (size_t)op->np() * op->nxz();
        auto nonneg = [](const float* p, size_t n) {
            for (size_t i = 0; i < n; i++)
                if (!(p[i] >= 0)) return false;
            return true;
        };
        if (w && !nonneg(w, ndata)) throw std::runtime_error("Data weights  must be non-negative");
        if (gmask && !nonneg(gmask, nmodel)) throw std::runtime_error("Gradient mask must be non-negative");
        std::unique_ptr<fwi2d_nlls> f(new fwi2d_nlls);
        f->op = op;
        f->ndata = ndata;
        f->nmodel = nmodel;
        f->integrate = (flags & FWI2D_NLLS_INTEGRATE) != 0;
        f->trace_norm = (flags & FWI2D_NLLS_TRACE_NORMALIZE) != 0;
        if (w) f->w.assign(w, w + ndata);
        if (gmask) f->gmask.assign(gmask, gmask + nmodel);
        if (filter) {
            for (int i = 0; i < nf; i++)
                if (!std::isfinite(filter[i])) throw std::runtime_error("the filter contains a non-finite value");
            f->filt.init(prepare_filter(filter, nf, filter_phase), filter_phase != FWI2D_FILTER_MINIMUM,
                         op->d.nt_data);
        }
        f->d.assign(d, d + ndata);
        f->process(f->d);
        if (flags & FWI2D_NLLS_NORMALIZE_OBJ) {
            double s = 0;
            for (double v : f->d) s += v * v;
            s *= op->d.dt_data;
            f->dnorm = s < 1e-16 ? 1.0 : s;
        }
        *out = f.release();
    });
}

int fwi2d_filter_prepare(const float* filter, int nf, int filter_phase, double* out, int* nout)
{
    return op_guard([&] {
        if (!filter || !out || !nout) throw std::runtime_error("null argument");
        if (nf < 1) throw std::runtime_error("the filter needs at least one sample");
        if (filter_phase < FWI2D_FILTER_DEFAULT || filter_phase > FWI2D_FILTER_MINIMUM)
            throw std::runtime_error("unknown filter phase");
        const std::vector<double> f = prepare_filter(filter, nf, filter_phase);
        std::copy(f.begin(), f.end(), out);
        *nout = (int)f.size();
    });
}

int fwi2d_filter_apply(const double* filter, int nf, int centered, int adjoint, long long ntr, int nt,
                       const double* in, double* out)
{
    return op_guard([&] {
        if (!filter || !in || !out) throw std::runtime_error("null argument");
        if (nf < 1 || nt < 1 || ntr < 0) throw std::runtime_error("invalid sizes");
        TraceFilter f;
        f.init(std::vector<double>(filter, filter + nf), centered != 0, nt);
        if (in < out + ntr * nt && out < in + ntr * nt) throw std::runtime_error("in and out overlap");
        parallel_traces<TraceFilter::Work>((size_t)(ntr + 1) / 2, [&](size_t p, TraceFilter::Work& w) {
            const long long x = 2 * (long long)p;
            const bool two = x + 1 < ntr;
            f.apply2(adjoint != 0, in + x * nt, out + x * nt, two ? in + (x + 1) * nt : nullptr,
                     two ? out + (x + 1) * nt : nullptr, w);
        });
    });
}

int fwi2d_nlls_evaluate(fwi2d_nlls* f, const float* m, double* func, float* g)
{
    return op_guard([&] {
        if (!f || !m) throw std::runtime_error("null argument");
        f->evaluate(m, func, g);
    });
}

const float* fwi2d_nlls_residual(fwi2d_nlls* f) { return f && !f->r.empty() ? f->r.data() : nullptr; }

double fwi2d_nlls_dnorm(fwi2d_nlls* f) { return f ? f->dnorm : 0.0; }

void fwi2d_nlls_destroy(fwi2d_nlls* f) { delete f; }

int fwi2d_nlls_objective(void* ctx, const float* m, double* f, float* g)
{
    return fwi2d_nlls_evaluate(static_cast<fwi2d_nlls*>(ctx), m, f, g);
}

}  // extern "C"
