// qp_envelope.h -- device state of a laser envelope (host side).
//
// No counterpart in the reference: there <a^2> is fixed for the whole run
// (src/System_3d.cpp:805-816 is its only writer) and chi is deposited
// (src/System_3d.cpp:396,526) but never read.  The envelope is the complex
// amplitude a of the laser vector potential Re[a exp(i k0 (x - t))] on the
// wake's grid (x toward the front, slice i at x = i dx), advanced in the time
// tau of the moving window by
//     lap_perp a + 2 d_tau (d_x + i k0) a = chi a.
// The scheme is stated in qp_envelope.hip.
#pragma once

#include "qp_internal.h"

namespace qp {

struct Envelope {
    qp_ctx_s *ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0, k0 = 0;
    Solver2d S;
    double *a = nullptr;     // 2 * nx * nn: real parts slice by slice, then the imaginary parts
    double *inc = nullptr;   // 2 slots x (re, im) x nn: a+ - a of the last two slices swept
    double *chi3 = nullptr;  // nx * nn: chi of every slice
    double *asq3 = nullptr;  // nx * nn: <a^2> = |a|^2 / 2
    double *cmid = nullptr;  // nx: midrange of chi per slice (set at each advance)
    int attached = 0;        // wakes reading this envelope
    hipGraphExec_t gexec = nullptr;  // one sweep, captured for dtau = g_dtau, g_iter chi corrections
    double g_dtau = 0.0;
    int g_iter = 0;
    hipStream_t g_stream = nullptr;  // stream the graph was captured on

    size_t nn() const { return (size_t)ny * nz; }
    size_t vol() const { return (size_t)nx * nn(); }
    hipStream_t stream() const { return ctx->stream; }
};

// One step of the whole envelope after a wake pass (midranges of chi, then the
// sweep: a captured graph when use_graph, else direct launches), enqueued on s.
int envelope_step(Envelope *e, double dt, int m, int use_graph, hipStream_t s);

}  // namespace qp

struct qp_envelope_s {
    qp::Envelope e;
};
