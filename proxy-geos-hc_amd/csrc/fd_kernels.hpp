// fd_kernels.hpp -- device data layout and kernel argument blocks of the FD
// (radius-R star stencil) step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phc {

// Field layout in HBM.  Point (i,j,k), i,j,k in [-R, n+R), lives at
//   origin + i*pitch_j + j*pitch_k + k.
// k is the contiguous axis (as in the reference's IDX3_l, FDTDmacros.hpp:14);
// each k-row is padded so that interior k = 0 starts on a 128-byte boundary
// (kofs = 32 floats, one L2 line) and every row starts 128-byte aligned.  Levels, vp^2dt^2
// and (for batched shots) per-shot arrays all use this layout.
struct FdLayout {
    int nx, ny, nz, R;
    int64_t pitch_k;  // floats per k-row
    int64_t pitch_j;  // floats per i-plane
    int64_t origin;   // offset of interior point (0,0,0)
    int64_t total;    // floats per array
};

inline FdLayout make_fd_layout(int nx, int ny, int nz, int R)
{
    FdLayout L;
    L.nx = nx; L.ny = ny; L.nz = nz; L.R = R;
    const int64_t kofs = 32;
    L.pitch_k = ((kofs + nz + R) + 31) / 32 * 32;
    // a row pitch that is a multiple of 2 KiB maps a tile's rows onto fewer
    // HBM channels (FD 448^3: 0.626 of the roofline at 512 floats, 0.653 at 576)
    if (L.pitch_k % 512 == 0) L.pitch_k += 64;
    L.pitch_j = L.pitch_k * (ny + 2 * R);
    L.origin = (int64_t)R * L.pitch_j + (int64_t)R * L.pitch_k + kofs;
    L.total = (int64_t)(nx + 2 * R) * L.pitch_j;
    return L;
}

__host__ __device__ inline int64_t fd_off(const FdLayout &L, int i, int j, int k)
{
    return L.origin + (int64_t)i * L.pitch_j + (int64_t)j * L.pitch_k + k;
}

constexpr int FD_MAX_R = 6;
constexpr int FD_MAX_SW = 64;  // widest sponge handled by the step kernels
constexpr int FD_TK = 64;  // k extent of a tile: one wavefront

struct FdStepArgs {
    const float *pn;    // current level (read, with neighbours)
    float *pnm1;        // previous level: read, then overwritten by the new level
    const float *vp;    // vp^2 dt^2
    FdLayout L;
    float cx[FD_MAX_R + 1], cy[FD_MAX_R + 1], cz[FD_MAX_R + 1];
    double coef0;
    float c0f;
    int sw;             // sponge width (0: no sponge)
    const float *fx, *fy, *fz;  // per-axis face damping values (1 outside the slab)
    float stab[FD_MAX_SW + 1];  // exp(alpha d^2), d = 0..sw (read with a uniform index)
    int64_t src_off;    // offset of the source point
    int damp_prev;      // previous level is read damped (deferred applySponge)
    int inj;            // add vp[src]*rhs_next at src after the update
    float rhs_next;
    float *trace;       // if set: *trace = new value at src before injection
    int tiles_k, tiles_j, chunk_i;
    // box of points this launch updates: i in [ilo, ihi) (marched in chunks
    // from ilo), j in [jlo, jhi), k in [klo, khi)
    int ilo, ihi, jlo, jhi, klo, khi;
    // PML (the marching kernel's PML variant): points outside the inner box
    // [px3, px4) x [py3, py4) x [pz3, pz4) take the pml3D update with
    // eta = pml_ex[i+1] + pml_ey[j+1] + pml_ez[k+1] (profiles over -1..n)
    const float *pml_ex, *pml_ey, *pml_ez;
    float *pml_phi;
    float pml_hd2[3];
    int px3, px4, py3, py4, pz3, pz4;
    // k-remainder slab: when the box's k extent ends in a partial 64-lane
    // tile of at most 32 lanes, the tiles cover k < krem only and the blocks
    // from nb_main on update the slab [krem, khi) one point per thread
    int nb_main;        // marching blocks (0: every block marches)
    int krem;
    // j-remainder slab: when the last TJ-row tile along j would hold at most
    // a quarter of its rows, the tiles cover j < jrem only and the same extra
    // blocks also update [klo, krem or khi) x [jrem, jhi) (0: none)
    int jrem;
    // per-XCD schedule (null: one block per (tile, chunk of chunk_i planes)
    // in XCD block order): hardware block b < nb_main marches tile column
    // sched[b].x (k tile + tiles_k * j tile) over planes [sched[b].y,
    // sched[b].z); an empty range does nothing
    const int4 *sched;
    // receiver sampling of the level this launch reads as current (the
    // previous step's result), by one extra block after the others:
    // rec_out[r * rec_stride + rec_step] = pn[rec_off[r]], or
    // rec_src[rec_step] (the value recorded before the fused injection) for a
    // receiver on the source point (0 receivers: no extra block)
    const int64_t *rec_off;
    float *rec_out;
    const float *rec_src;
    int rec_n, rec_stride, rec_step;
};

} // namespace phc
