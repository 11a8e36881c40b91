// The reference's l-BFGS solver with its regular Wolfe line search (lbfgs::run,
// updateSY, computeHg: src/cpp/nlsolver.cpp:1207-1415; regular_wolfe::lineSearch and
// zoom: nlsolver.cpp:184-421; quadratic / cubic step interpolation: nlsolver.cpp:9-35),
// restated on plain arrays over an objective callback, so that the host loop of an FWI
// run drives the batched objective (fwi2d_nlls_*) without the reference's vecReg
// machinery.  Each evaluation returns the functional and the gradient together, as
// nlls_fwi_eco's res() does (optimization.hpp:592-909).  Vectors are single precision
// like the reference's data_t; dot products and norms accumulate in double precision.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "fwi2d_hip.h"

namespace fwi2d {
extern thread_local std::string g_last_error;
}

namespace {

const double kZero = 1e-10;  // nlsolver.cpp:6
const double kStol = 1e-3;   // nlsolver.cpp:7
const double kNaN = std::numeric_limits<double>::quiet_NaN();

double dot(const std::vector<float>& a, const std::vector<float>& b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += (double)a[i] * b[i];
    return s;
}

// minimiser of the quadratic through (0, f, df) and (stp0, f0) (nlsolver.cpp:9-15)
double quadratic(double f, double df, double stp0, double f0)
{
    double ans = (f0 <= f + df * stp0) ? stp0 * 1.1 : -df * stp0 * stp0 / (2 * (f0 - f - df * stp0));
    if (ans / stp0 < kStol || std::abs(ans - stp0) / stp0 < kStol) ans = stp0 / 2.0;
    return ans;
}

// minimiser of the cubic through (0, f, df), (stp0, f0), (stp1, f1) (nlsolver.cpp:17-35)
double cubic(double f, double df, double stp0, double f0, double stp1, double f1)
{
    const double x1 = f1 - f - df * stp1, x2 = f0 - f - df * stp0;
    const double x3 = stp0 * stp0 * stp1 * stp1 * (stp1 - stp0);
    const double a = (stp0 * stp0 * x1 - stp1 * stp1 * x2) / x3;
    const double b = (-stp0 * stp0 * stp0 * x1 + stp1 * stp1 * stp1 * x2) / x3;
    const double delta = b * b - 3 * a * df;
    const double step = f1 < f0 ? stp1 : stp0, func = f1 < f0 ? f1 : f0;
    double ans = (std::abs(a) < kZero || delta < 0) ? quadratic(f, df, step, func) : (-b + std::sqrt(delta)) / (3 * a);
    if (ans / step < kStol || std::abs(step - ans) / step < kStol) ans = step / 2.0;
    else if (std::abs(stp0 + stp1 - step - ans) / (stp0 + stp1 - step) < kStol) ans = (stp0 + stp1 - step) / 2.0;
    return ans;
}

struct Problem {
    fwi2d_objective_fn fn;
    void* ctx;
    std::vector<float> m, g;  // current model and the gradient of the last evaluation
    double f = 0;
    int feval = 0, geval = 0;  // counted as lbfgs::run counts them (nlsolver.cpp:1239, 1268; 242, 275)
    fwi2d_lbfgs_report* report = nullptr;
    int iter = 0, trial = 0;  // iteration and trial of the line search in progress
    // res() + getFunc(): the objective returns the gradient with the functional, as
    // nlls_fwi_eco's compute_res_and_grad does; it is counted when the search reads it
    void eval()
    {
        if (fn(ctx, m.data(), &f, g.data()) != 0)
            throw std::runtime_error("objective evaluation failed: " + fwi2d::g_last_error);
        if (f != f) throw std::runtime_error("the objective function is NaN");
        feval++;
    }
    // one trial of a line search: its step length and functional, and the gradient norm
    // if the search reads the gradient (the reference's verbose trial block)
    void record(double alpha, double gnorm)
    {
        if (!report) return;
        const int k = report->ntrials++;
        if (report->trials && k < report->trials_cap) {
            double* row = report->trials + 5 * (size_t)k;
            row[0] = iter;
            row[1] = trial;
            row[2] = alpha;
            row[3] = f;
            row[4] = gnorm;
        }
    }
};

// regular_wolfe (nlsolver.hpp:85-155): the state carried across iterations
struct RegularWolfe {
    double c1, c2, a0, a1, stp_max;
    int version;
    double stp0 = 0, stp1 = 0, stp = 0, f = 0, df = 0, f0 = 0, f1 = 0;

    // phi(alpha) = f(m0 + alpha p): the model is set and evaluated
    double trial(Problem& P, const std::vector<float>& m0, const std::vector<float>& p, double alpha)
    {
        for (size_t i = 0; i < P.m.size(); i++) P.m[i] = (float)(m0[i] + alpha * p[i]);
        P.eval();
        return P.f;
    }

    // lineSearch (nlsolver.cpp:184-312)
    bool search(Problem& P, const std::vector<float>& m0, const std::vector<float>& p, int iter, int max_trial,
                double& gnorm)
    {
        const double df0 = df;
        f = P.f;
        df = dot(P.g, p);
        stp0 = 0;
        double alpha;
        if (iter == 0) {
            f0 = f;
            stp1 = a0 / gnorm;
            if (a1 > 0) stp1 = std::max(stp1, a1 * std::sqrt(dot(P.m, P.m)) / gnorm);
        } else {
            stp1 = stp * df0 / df;
            if (version == 1) {
                stp1 = 2 * (f - f0) / df;
                stp1 = std::min(1.0, 1.00001 * stp1);
            }
        }
        alpha = stp1;
        int n_trial = 0;
        while (true) {
            if (n_trial >= max_trial) return false;
            const double fa = trial(P, m0, p, alpha);
            n_trial++;
            P.trial = n_trial;
            f1 = fa;
            if (fa <= kZero) {
                P.record(alpha, kNaN);
                stp = alpha;
                f0 = f;
                f = fa;
                return false;
            }
            if (f1 > f + c1 * alpha * df || (f1 >= f0 && n_trial > 1)) {
                P.record(alpha, kNaN);
                return zoom(false, P, m0, p, n_trial, max_trial, gnorm);
            }
            P.geval++;  // phi'(alpha): the gradient is read
            gnorm = std::sqrt(dot(P.g, P.g));
            P.record(alpha, gnorm);
            const double df1 = dot(P.g, p);
            if (df1 >= c2 * df) {  // curvature condition
                stp = alpha;
                f0 = f;
                f = fa;
                return true;
            }
            if (df1 >= 0) return zoom(true, P, m0, p, n_trial, max_trial, gnorm);
            stp0 = stp1;  // extrapolate towards stp_max
            stp1 = std::min(quadratic(f, df, stp1, f1), stp_max);
            alpha = stp1;
            f0 = f1;
        }
    }

    // zoom (nlsolver.cpp:314-421)
    bool zoom(bool reverse, Problem& P, const std::vector<float>& m0, const std::vector<float>& p, int& n_trial,
              int max_trial, double& gnorm)
    {
        double alo = stp0, ahi = stp1, flo = f0, fhi = f1;
        if (reverse) {
            alo = stp1;
            ahi = stp0;
            flo = f1;
            fhi = f0;
        }
        while (true) {
            if (n_trial >= max_trial) return false;
            double alpha;
            if (alo == 0) alpha = quadratic(f, df, ahi, fhi);
            else if (ahi == 0) alpha = quadratic(f, df, alo, flo);
            else alpha = cubic(f, df, alo, flo, ahi, fhi);
            const double fa = trial(P, m0, p, alpha);
            n_trial++;
            P.trial = n_trial;
            f0 = f1;
            f1 = fa;
            if (fa <= kZero) {
                P.record(alpha, kNaN);
                stp = alpha;
                f0 = f;
                f = fa;
                return false;
            }
            if (f1 > f + c1 * alpha * df || f1 >= flo) {
                P.record(alpha, kNaN);
                ahi = alpha;
                fhi = f1;
            } else {
                P.geval++;
                gnorm = std::sqrt(dot(P.g, P.g));
                P.record(alpha, gnorm);
                const double df1 = dot(P.g, p);
                if (df1 >= c2 * df) {
                    stp = alpha;
                    f0 = f;
                    f = fa;
                    return true;
                }
                if (df1 * (ahi - alo) >= 0) {
                    ahi = alo;
                    fhi = flo;
                }
                alo = alpha;
                flo = f1;
            }
        }
    }
};

}  // namespace

extern "C" {

int fwi2d_lbfgs(fwi2d_objective_fn fn, void* ctx, long long n, float* m, const fwi2d_lbfgs_options* opt,
                double* func_history, int* iterations, fwi2d_lbfgs_report* report)
{
    try {
        if (report) {
            if (report->trials_cap < 0) throw std::runtime_error("negative trials_cap");
            report->feval = report->geval = report->ntrials = 0;
        }
        if (!fn || !m || !opt || n < 1) throw std::runtime_error("null argument");
        const fwi2d_lbfgs_options& o = *opt;
        if (!(o.c1 > 0 && o.c1 < 1 && o.c2 > o.c1 && o.c2 < 1))
            throw std::runtime_error("The constants c1, c2 must satisfy 0<c1<c2<1");
        if (o.niter < 0 || o.max_trial < 1 || o.memory < 1) throw std::runtime_error("invalid solver options");
        std::vector<float> H0((size_t)n, 1.0f);
        if (o.H0) {
            for (long long i = 0; i < n; i++)
                if (!(o.H0[i] > 0)) throw std::runtime_error("Initial diagonal inverse Hessian must be definite positive.");
            H0.assign(o.H0, o.H0 + n);
        }
        Problem P{fn, ctx, std::vector<float>(m, m + n), std::vector<float>((size_t)n, 0.f)};
        P.report = report;
        RegularWolfe ls{o.c1, o.c2, o.a0, o.a1, o.step_max, o.version};
        std::vector<double> func;
        auto finish = [&](int k) {
            std::memcpy(m, P.m.data(), (size_t)n * sizeof(float));
            if (func_history)
                for (size_t i = 0; i < func.size(); i++) func_history[i] = func[i];
            if (iterations) *iterations = k;
            if (report) {
                report->feval = P.feval;
                report->geval = P.geval;
            }
        };
        // lbfgs::run (nlsolver.cpp:1207-1346)
        P.eval();
        P.geval++;  // the gradient of the starting model (nlsolver.cpp:1266-1269)
        func.push_back(P.f);
        if (func[0] <= kZero) {
            finish(0);
            fwi2d::g_last_error.clear();
            return 0;
        }
        const int mem = o.memory;
        std::vector<std::vector<float>> s(mem, std::vector<float>((size_t)n, 0.f)), y = s;
        std::vector<double> rho(mem, 0.0);
        std::vector<float> m0 = P.m, g0((size_t)n), p((size_t)n), q((size_t)n);
        int k = 0;
        double rate = 1, gnorm = std::sqrt(dot(P.g, P.g));
        bool success = true;
        while (k < o.niter && rate > o.threshold && gnorm > kZero && P.f > kZero && success) {
            // updateSY (nlsolver.cpp:1348-1386): the newest pair last, the oldest dropped
            if (k > 0) {
                int slot = k - 1;
                if (k > mem) {
                    std::rotate(s.begin(), s.begin() + 1, s.end());
                    std::rotate(y.begin(), y.begin() + 1, y.end());
                    std::rotate(rho.begin(), rho.begin() + 1, rho.end());
                    slot = mem - 1;
                }
                for (long long i = 0; i < n; i++) {
                    s[slot][i] = P.m[i] - m0[i];
                    y[slot][i] = P.g[i] - g0[i];
                }
                rho[slot] = dot(s[slot], y[slot]);
            }
            // computeHg (nlsolver.cpp:1388-1415): the two-loop recursion, p = -H g
            const int mm = std::min(k, mem);
            std::vector<double> a((size_t)mm, 0.0);
            q = P.g;
            double gamma = 1.0;
            for (int i = mm - 1; i >= 0; i--) {
                a[i] = dot(s[i], q) / rho[i];
                for (long long j = 0; j < n; j++) q[j] = (float)(q[j] - a[i] * y[i][j]);
            }
            if (mm > 0) gamma = rho[mm - 1] / dot(y[mm - 1], y[mm - 1]);
            for (long long j = 0; j < n; j++) p[j] = (float)(gamma * H0[j] * q[j]);
            for (int i = 0; i < mm; i++) {
                const double b = dot(y[i], p) / rho[i];
                for (long long j = 0; j < n; j++) p[j] = (float)(p[j] + (a[i] - b) * s[i][j]);
            }
            for (auto& v : p) v = -v;
            if (dot(p, P.g) >= 0) break;  // not a descent direction
            if (k > 0) m0 = P.m;
            g0 = P.g;
            P.iter = k;
            success = ls.search(P, m0, p, k, o.max_trial, gnorm);
            if (!success) break;
            func.push_back(P.f);
            if (k > 0) rate = (func[k] - func[k + 1]) / func[k];
            k++;
        }
        finish(k);
        fwi2d::g_last_error.clear();
        return 0;
    } catch (const std::exception& ex) {
        fwi2d::g_last_error = ex.what();
        return 1;
    }
}

}  // extern "C"
