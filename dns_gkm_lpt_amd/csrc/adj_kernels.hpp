// adj_kernels.hpp -- tangent-linear and adjoint of the fixed-dt FP64 step.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "gkm_math.hpp"

namespace gkm {

// device scratch of one tangent-linear / adjoint step (face arrays of nc^3 each)
struct AdjScratch {
    double* WS;     // 10: WENO states Wl, Wr of the current direction
    double* WSt;    // 10: their tangents
    double* Ft[3];  // 5 each: flux tangents per direction
    double* Xbar;   // 50: adjoints of the flux inputs (WL, WR, Wl, Wr, 6 slopes)
    double* WSbar;  // 10: adjoints of the WENO states
    double* CL;     // 5: adjoint of the face's left cell
    double* CR;     // 5: adjoint of the face's right cell
    double* Sbar;   // 30: adjoints of the WENO stencil cells
};
size_t adj_scratch_doubles(long plane);
void adj_scratch_carve(double* base, long plane, AdjScratch& s);

// Wt_out = dStep(W) . Wt  (state and tangents: 5 planes of nc^3, periodic)
hipError_t launch_tl_step(const double* W, const double* Wt, long plane, int nc, const Gas<double>& g, double dt,
                          double dx, const AdjScratch& s, double* Wt_out, hipStream_t st);
// lam_out = dStep(W)^T . lam
hipError_t launch_adj_step(const double* W, const double* lam, long plane, int nc, const Gas<double>& g, double dt,
                           double dx, const AdjScratch& s, double* lam_out, hipStream_t st);

}  // namespace gkm
