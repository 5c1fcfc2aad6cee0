// lb.hpp -- learned thin absorbing boundary of the FD step: shared constants,
// parameter layout and the hooks the FD step calls.
//
// No reference counterpart.  The reference absorbs outgoing waves with a
// sponge (src/fd/utilsFDTD/FDTDinit.hpp:319-396, FDTDkernels.hpp:108-118) or a
// PML (FDTDkernels.hpp:36-104).  Here the grid ends LB_W points outside the
// region of interest, with no damping.  At a point of face f's layer (depth d =
// 0..LB_W-1 from the grid edge) the radius-4 update reads 4-d points beyond
// the edge along f's normal; the grid's halo holds a constant there.  The
// missing term
//      E = sum_{l=d+1..4} c_l (p(P + l n_out) - h)
// (c_l the stencil coefficients, h the halo value) is what a transparent
// boundary would add.  Face f's MLP predicts it from what the grid holds:
//      x = ((pn - h)/s along the inward normal, offsets -4..4;
//           (pnm1 - h)/s, same offsets; sqrt(vp^2dt^2 c_1)),
//      s = max |inputs - h|,
//      E ~= c_1 s y,   y = w3 . relu(W2 relu(W1 x + b1[d]) + b2) + b3,
// and the new value at the point gains vp^2dt^2 * c_1 * s * y.  Scaling by s
// makes the prediction proportional to the local amplitude (the wave
// equation is linear), so one network serves every source strength.
//
// At run time y is clamped to +-ymax(d) = sum_{l=d+1..4} |c_l / c_1|: the
// largest |E| / (c_1 s) a field no larger beyond the edge than inside the
// stencil window can give.  Without the bound a layer trained with input
// noise slowly feeds energy back in long runs (its network, evaluated on the
// low-amplitude remainder it meets late in a run, is outside what it was
// trained on); with it an 8000-step run stays below the bare grid's
// amplitude while the 2000-step traces are unchanged to 1%.  The bound does
// not rescue a layer trained without input noise.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fd_kernels.hpp"
#include "phc.h"

namespace phc {

constexpr int LB_W = 4;       // layer width (points)
constexpr int LB_R = 4;       // stencil half width the layer is built for
constexpr int LB_NIN = 19;    // 9 pn + 9 pnm1 + velocity feature
constexpr int LB_KIN = 24;    // inputs padded to 3 MFMA k-chunks of 8
constexpr int LB_H = 32;      // hidden width (both layers)
constexpr int LB_FACES = 6;   // face = 2*axis + side; axis 0 = i, 1 = j, 2 = k; side 0 = low
// parameters of one face, in this order
constexpr int LB_OFF_W1 = 0;                          // [LB_H][LB_NIN]
constexpr int LB_OFF_B1 = LB_OFF_W1 + LB_H * LB_NIN;  // [LB_W][LB_H] (per layer depth)
constexpr int LB_OFF_W2 = LB_OFF_B1 + LB_W * LB_H;    // [LB_H][LB_H]
constexpr int LB_OFF_B2 = LB_OFF_W2 + LB_H * LB_H;    // [LB_H]
constexpr int LB_OFF_W3 = LB_OFF_B2 + LB_H;           // [LB_H]
constexpr int LB_OFF_B3 = LB_OFF_W3 + LB_H;           // [1]
constexpr int LB_NP = LB_OFF_B3 + 1;                  // 1825 per face
constexpr int LB_NSTRIDE = 22;                        // floats per training sample
constexpr float LB_S_TINY = 1e-30f;                   // amplitudes at or below: quiet point

// A training sample: x[LB_NIN], target, depth (as float), unused.
// Device-side inference state attached to an FD handle.
struct LbInfer {
    const uint16_t *w;   // bf16 weights, LB_FACES x LB_NP
    float *corr;         // per-face compact corrections, see lb_corr_offset
    int64_t corr_off[LB_FACES + 1];
    float halo;          // value held by the halo (h)
    float c1;            // c_1 of the stencil
    float chat[LB_R + 1];  // c_l / c_1
    float ldamp[LB_W];   // new value at layer depth d multiplied by ldamp[d] (per face holding it)
    // k-faces (f = 4, 5; their rows end at the grid's contiguous axis, so a
    // row end costs a whole line): vp c1 and sqrt(vp c1) at the layer points,
    // kvcs[((f - 4) * 2 + which) * LB_W + d][u nv + v], computed once; and two
    // stores of the LB_W + LB_R in-grid planes of a level, kslab[s][((f - 4)
    // * (LB_W + LB_R) + q) * npts + u nv + v], which a correction pass can
    // write (its current level) and the next one read (as its previous).
    const float *kvcs;
    float *kslab[2];
    int npts_k;          // u nv + v range of a k-face (nx ny)
};

// Bound of the network's output at layer depth d (see above), in float.
__host__ __device__ inline float lb_ymax(const float chat[LB_R + 1], int d)
{
    float m = 0.f;
    for (int l = d + 1; l <= LB_R; ++l) m += chat[l] < 0.f ? -chat[l] : chat[l];
    return m;
}

// Launch the correction pass (before the step: reads the current and previous
// levels) and the apply pass (after the step: adds the corrections to the new
// level at layer points).
// With kslab_in the k-faces' previous-level planes come from that store
// (written by the previous pass) instead of pnm1; with kslab_out the pass
// writes its current level's k-face planes there.
hipError_t lb_enqueue_correct(const LbInfer &lb, const FdLayout &L, const float *pn, const float *pnm1,
                              const float *vp, hipStream_t stream, const float *kslab_in = nullptr,
                              float *kslab_out = nullptr);
// vp c1 and sqrt(vp c1) at the k-faces' layer points into lb.kvcs.
hipError_t lb_enqueue_kface_vp(const LbInfer &lb, const FdLayout &L, const float *vp, float *kvcs,
                               hipStream_t stream);
hipError_t lb_enqueue_apply(const LbInfer &lb, const FdLayout &L, float *pnew, hipStream_t stream);
// Size of the compact correction store for a grid, filling corr_off.
int64_t lb_corr_layout(const FdLayout &L, int64_t *corr_off);
// Copy a model's bf16 parameters (LB_FACES x LB_NP) to device memory.
phc_status lb_export_bf16(const phc_lb *h, uint16_t *dev_dst, hipStream_t stream);

// What the training-data collection reads from an FD handle (fd.hip).
struct FdFieldView {
    FdLayout L;
    const float *lev[2];
    const float *vp;
    int sw;
    float c1;              // c_1 of the stencil (equal on the three axes)
    float chat[LB_R + 1];  // c_l / c_1
    double coef0;          // centre coefficient
    hipStream_t stream;
};
phc_status fd_field_view(phc_fd *h, FdFieldView *v);

} // namespace phc
