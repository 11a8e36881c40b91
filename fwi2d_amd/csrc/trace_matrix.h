// Trace processing of the operator level on the device.
//
// Every trace goes through the same linear map in time: the data <-> propagation
// resampling (sinc_resampler / linear_resampler, operator.cpp:760-856), the time
// derivative of particle-velocity data (Dt, misc.cpp:71-95), the dt ratio, and for the
// adjoint sources the inverse trapezoid weights (applyHt, misc.cpp:45-69) and the time
// reversal (we_op.cpp:1112-1127).  The host composes that map into one row-compressed
// matrix in double precision; the device applies it to all traces at once.
#pragma once
#include <stdexcept>
#include <vector>

#include "fwi2d_hip.h"

namespace fwi2d {

// out[x][i] = sum over k in [ptr[i], ptr[i+1]) of w[k] * in[x][col[k]], for every trace x
struct TraceMatrix {
    int nrow = 0, ncol = 0;
    std::vector<int> ptr, col;
    std::vector<double> w;
};

// the matrix has ncol_expected columns / nrow_expected rows (when >= 0) and a valid
// row-compressed structure
inline void check_trace_matrix(const TraceMatrix& m, int ncol_expected, int nrow_expected)
{
    if ((ncol_expected >= 0 && m.ncol != ncol_expected) || (nrow_expected >= 0 && m.nrow != nrow_expected) ||
        m.nrow < 0 || m.ncol < 1 || (int)m.ptr.size() != m.nrow + 1 || m.ptr[0] != 0 ||
        (size_t)m.ptr[m.nrow] != m.col.size() || m.col.size() != m.w.size())
        throw std::runtime_error("trace matrix does not match the traces");
    for (int i = 0; i < m.nrow; i++)
        if (m.ptr[i + 1] < m.ptr[i]) throw std::runtime_error("trace matrix rows out of order");
    for (int c : m.col)
        if (c < 0 || c >= m.ncol) throw std::runtime_error("trace matrix column out of range");
}

// Forward run of every shot (sources at the propagation step, host memory) whose receiver
// traces [2 nr][nt] are mapped by m (ncol = nt) on the device into out [2 nr][m.nrow] (host);
// DAS channels (a gauge length set with the acquisition) then go to strain on the device
// (dipole_to_strain, misc.cpp:18-43).
void engine_forward_traces(fwi2d_engine* e, const float* src, bool keep, const TraceMatrix& m, float* out);

// Adjoint run whose adjoint sources are m applied on the device to data [2 nr][m.ncol]
// (host, m.nrow = nt; DAS channels: dipole_to_strain's adjoint first); the raw gradient
// [np][nx][nz] goes to grad (host).
void engine_adjoint_traces(fwi2d_engine* e, const float* data, const TraceMatrix& m, float* grad);

// The same for the acoustic engine: one trace component, receiver traces [nr][nt]
// mapped into out [nr][m.nrow]; the adjoint's raw gradient is [2][nx][nz] (K, 1/rho).
void acoustic_forward_traces(fwi2d_acoustic* e, const float* src, bool keep, const TraceMatrix& m, float* out);
void acoustic_adjoint_traces(fwi2d_acoustic* e, const float* data, const TraceMatrix& m, float* grad);

}  // namespace fwi2d
