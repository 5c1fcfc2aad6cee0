// fd.hip -- FD step kernels (radius-R star stencil + leapfrog + sponge +
// source injection) and the phc_fd_* C ABI.
//
// Reference hot path (src/fd/utilsFDTD/):
//   FDTDKernel::addRHS           FDTDkernels.hpp:121-128
//   FDTDKernel::inner3D          FDTDkernels.hpp:8-34
//   FDTDKernel::applySponge      FDTDkernels.hpp:108-118
//   FDTDKernel::computeOneStepSB FDTDkernels.hpp:75-84
//   time loop                    FDTDmain.cpp:55-81, output FDTDutils.hpp:193-232
//
// One launch does inner3D + applySponge for every interior point.  A block
// owns a 64(k) x TJ(j) column of the grid and marches along i over chunk_i
// planes.  Each thread keeps the 2R+1 values of pn along i for its NJ points
// in registers; the current i-plane of the tile plus its j/k halo is staged in
// LDS for the in-plane neighbours.  HBM traffic per point: read pn, pnm1, vp,
// write the new level (16 B); halo re-reads are served from L2/MALL.
//
// The reference damps the current level in place after the stencil
// (applySponge's second product).  Here that product is applied when the
// level is next read as "previous" (damp_prev), which gives the same float
// values without a second pass over the grid.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fd_kernels.hpp"
#include "fd_tile.hpp"
#include "lb.hpp"
#include "march.hpp"
#include "phc_common.hpp"
#include "xcd_sched.hpp"

namespace phc {

// ----------------------------------------------------------------- kernels

// Tile column and plane range of a marching block: from the per-XCD table,
// or tiles in XCD block order with chunks of chunk_i planes.
__device__ __forceinline__ void fd_block_range(const FdStepArgs &a, int &kt, int &jt, int &i0, int &i1)
{
    int w;
    if (a.sched) {
        const int4 e = a.sched[blockIdx.x];
        w = e.x;
        i0 = e.y;
        i1 = e.z;
    } else {
        const int b = xcd_block(blockIdx.x, a.nb_main ? a.nb_main : gridDim.x - (a.rec_n > 0 ? 1 : 0));
        const int nt = a.tiles_k * a.tiles_j, ic = b / nt;
        w = b - ic * nt;
        i0 = a.ilo + ic * a.chunk_i;
        i1 = min(i0 + a.chunk_i, a.ihi);
    }
    kt = w % a.tiles_k;
    jt = w / a.tiles_k;
}

// The extra block's receiver sampling (FdStepArgs::rec_n).
__device__ __forceinline__ void fd_sample_prev(const FdStepArgs &a)
{
    for (int r = (int)(threadIdx.y * blockDim.x + threadIdx.x); r < a.rec_n; r += (int)(blockDim.x * blockDim.y)) {
        const int64_t o = a.rec_off[r];
        a.rec_out[(int64_t)r * a.rec_stride + a.rec_step] = (o == a.src_off) ? a.rec_src[a.rec_step] : a.pn[o];
    }
}

template <int R, int TY, int NJ, int ARITH, typename IT>
__global__ void __launch_bounds__(FD_TK *TY)
fd_step_kernel(FdStepArgs a)
{
    using SH = FdTileShape<R, TY, NJ>;
    __shared__ float tile[2][SH::LH][SH::LW];
    if (a.rec_n > 0 && blockIdx.x == gridDim.x - 1) {
        fd_sample_prev(a);
        return;
    }
    if (a.nb_main && (int)blockIdx.x >= a.nb_main) {
        fd_rem_points<R, ARITH, IT>(a, a, (int64_t)(blockIdx.x - a.nb_main) * (FD_TK * TY), FD_TK * TY);
        return;
    }
    int kt, jt, i0, i1;
    fd_block_range(a, kt, jt, i0, i1);
    if (i0 >= i1) return;  // padding of a shorter XCD list (whole block)
    fd_tile<R, TY, NJ, ARITH, IT>(a, a, kt, jt, i0, i1, tile);
}

// computeOneStepPML (FDTDkernels.hpp:87-104) in one launch: the march over
// the whole grid, pml3D (FDTDkernels.hpp:36-73) on the slab points, inner3D
// inside.
template <int R, int TY, int NJ, int ARITH, typename IT>
__global__ void __launch_bounds__(FD_TK *TY)
fd_step_pml_kernel(FdStepArgs a)
{
    using SH = FdTileShape<R, TY, NJ>;
    __shared__ float tile[2][SH::LH][SH::LW];
    if (a.rec_n > 0 && blockIdx.x == gridDim.x - 1) {
        fd_sample_prev(a);
        return;
    }
    if (a.nb_main && (int)blockIdx.x >= a.nb_main) {
        fd_rem_points<R, ARITH, IT, true>(a, a, (int64_t)(blockIdx.x - a.nb_main) * (FD_TK * TY), FD_TK * TY);
        return;
    }
    int kt, jt, i0, i1;
    fd_block_range(a, kt, jt, i0, i1);
    if (i0 >= i1) return;  // padding of a shorter XCD list (whole block)
    fd_tile<R, TY, NJ, ARITH, IT, true>(a, a, kt, jt, i0, i1, tile);
}

// addRHS (FDTDkernels.hpp:121-128): one point.
__global__ void fd_add_rhs_kernel(float *p, const float *vp, int64_t off, const float *rhs, int it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] = add_nc(p[off], mul_nc(vp[off], rhs[it]));
}

// addRHS with the sample value handed over by the caller.
__global__ void fd_add_value_kernel(float *p, const float *vp, int64_t off, float rhs)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) p[off] = add_nc(p[off], mul_nc(vp[off], rhs));
}

// Deferred sponge product of a level: p *= s on every interior point.
__global__ void fd_damp_kernel(float *p, FdLayout L, int sw, const float *fx, const float *fy,
                               const float *fz)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int i = blockIdx.z;
    if (k >= L.nz) return;
    const bool inI = (i >= sw) && (i < L.nx - sw);
    const bool inJ = (j >= sw) && (j < L.ny - sw);
    const bool inK = (k >= sw) && (k < L.nz - sw);
    float s = 1.f;
    if (inJ && inK && !inI) s = fx[i];
    if (inK && inI && !inJ) s = fy[j];
    if (inJ && inI && !inK) s = fz[k];
    const int64_t o = fd_off(L, i, j, k);
    p[o] = mul_nc(p[o], s);
}

// Receiver sampling after a step: out[r*stride + step] = level[off[r]].  A
// receiver on the source point reads the value recorded before the fused
// injection of the next step.
__global__ void fd_sample_kernel(const float *p, const int64_t *offs, int nrec, float *out,
                                 int stride, int step, int64_t src_off, const float *src_trace)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int64_t o = offs[r];
    out[(int64_t)r * stride + step] = (o == src_off) ? src_trace[step] : p[o];
}

// ------------------------------------------------------------ dispatching

using FdKernelFn = void (*)(FdStepArgs);

template <int R, int TY, int NJ>
static FdKernelFn pick_arith(int arith, bool idx32, bool pml)
{
    if (pml) {
        if (arith == PHC_FD_ARITH_REF)
            return idx32 ? fd_step_pml_kernel<R, TY, NJ, PHC_FD_ARITH_REF, uint32_t>
                         : fd_step_pml_kernel<R, TY, NJ, PHC_FD_ARITH_REF, int64_t>;
        return idx32 ? fd_step_pml_kernel<R, TY, NJ, PHC_FD_ARITH_FP32, uint32_t>
                     : fd_step_pml_kernel<R, TY, NJ, PHC_FD_ARITH_FP32, int64_t>;
    }
    if (arith == PHC_FD_ARITH_REF)
        return idx32 ? fd_step_kernel<R, TY, NJ, PHC_FD_ARITH_REF, uint32_t>
                     : fd_step_kernel<R, TY, NJ, PHC_FD_ARITH_REF, int64_t>;
    return idx32 ? fd_step_kernel<R, TY, NJ, PHC_FD_ARITH_FP32, uint32_t>
                 : fd_step_kernel<R, TY, NJ, PHC_FD_ARITH_FP32, int64_t>;
}

// Block shapes: TY wavefronts per block, NJ rows per thread (tile_j = TY*NJ).
// Radius 4 (the 8th-order operator and order-4 SEM) gets the full set; the
// other radii one shape.
template <int R>
static FdKernelFn pick_shape(int ty, int nj, int arith, bool idx32, bool pml)
{
    if (ty == 4 && nj == 2) return pick_arith<R, 4, 2>(arith, idx32, pml);
    if constexpr (R == 4) {
        if (ty == 4 && nj == 4) return pick_arith<R, 4, 4>(arith, idx32, pml);
        if (ty == 8 && nj == 1) return pick_arith<R, 8, 1>(arith, idx32, pml);
        if (ty == 8 && nj == 2) return pick_arith<R, 8, 2>(arith, idx32, pml);
        if (ty == 16 && nj == 1) return pick_arith<R, 16, 1>(arith, idx32, pml);
        if (ty == 16 && nj == 2) return pick_arith<R, 16, 2>(arith, idx32, pml);
    }
    return nullptr;
}

// Kernel for half width R and block shape, arithmetic mode, 32-bit offsets
// when every array is < 4 GiB, with or without the PML slabs.
FdKernelFn fd_pick_kernel(int R, int ty, int nj, int arith, bool idx32, bool pml)
{
    switch (R) {
    case 1: return pick_shape<1>(ty, nj, arith, idx32, pml);
    case 2: return pick_shape<2>(ty, nj, arith, idx32, pml);
    case 3: return pick_shape<3>(ty, nj, arith, idx32, pml);
    case 4: return pick_shape<4>(ty, nj, arith, idx32, pml);
    case 5: return pick_shape<5>(ty, nj, arith, idx32, pml);
    case 6: return pick_shape<6>(ty, nj, arith, idx32, pml);
    default: return nullptr;
    }
}

// Per-axis face damping values (FDTDinit.hpp:343-395): for the low face
// exp(alpha (W-i)^2), for the high face exp(alpha (W-(n-i))^2); the high-face
// loop runs second in the reference, so it wins where the two overlap.
void fd_axis_damping(int n, int W, float alpha, std::vector<float> &f)
{
    f.assign(n, 1.f);
    for (int i = 0; i < n; ++i) {
        int d = -1;
        if (i >= n - W) d = W - (n - i);
        else if (i < W) d = W - i;
        if (d >= 0) f[i] = (float)std::exp((double)alpha * std::pow((double)d, 2.0));
    }
}

} // namespace phc

// ====================================================================== API
using namespace phc;

struct phc_fd {
    int nx, ny, nz, R;  // R: the largest half width (the stored halo on every axis)
    int lh[3];          // per-axis half widths lx, ly, lz (the reference's halos)
    int xs, ys, zs;
    int arith;
    int sw;
    float alpha;
    float cx[FD_MAX_R + 1], cy[FD_MAX_R + 1], cz[FD_MAX_R + 1];
    double coef0;
    int nsamples;
    FdLayout L;
    DevBuf lev[2], vp, fx, fy, fz, rhs, trace, rec, recoff;
    std::vector<float> hfx, hfy, hfz, hvp, hrhs;
    int pending[2] = {0, 0};
    hipStream_t stream = nullptr;
    int ty = 4, nj = 2, chunk_i = 0, index64 = 0;
    int krem_mode = 0;  // phc_fd_params.k_remainder (as requested)
    int krem_eff = 0;   // the mode the launches use (the autotuner may pick -1 for 0)
    // chunk_i < 0: the per-XCD schedule table, rebuilt when its key changes
    int cus = 256;
    DevBuf sched;
    int sched_nb = 0;
    long long sched_key = -1;
    int nrec = 0;
    std::vector<int64_t> rec_offs;
    int last_n = 0, last_nrec = 0;
    int timing = 0;
    double last_ms = 0.0;
    // PML (phc_fd_set_pml): slab widths, eta profiles, phi, 1/(4 d^2)
    int pml_on = 0;
    int pml_nd[3] = {0, 0, 0};
    DevBuf pml_ex, pml_ey, pml_ez, pml_phi;
    float pml_hd2[3] = {0.f, 0.f, 0.f};
    // learned thin boundary (phc_fd_set_learned_boundary)
    int lb_on = 0;
    LbInfer lbi;
    DevBuf lb_w, lb_corr, lb_kvcs, lb_kslab;
    // cached graph of the last phc_fd_run
    hipGraphExec_t gexec = nullptr;
    int g_it0 = -1, g_n = -1, g_cur = -1, g_damp = -1, g_nrec = -1;
};

static phc_status fd_validate(const phc_fd_params *p)
{
    if (!p) return fail(PHC_ERR_INVALID, "params is NULL");
    if (p->nx < 1 || p->ny < 1 || p->nz < 1)
        return fail(PHC_ERR_INVALID, "grid %d x %d x %d is empty", p->nx, p->ny, p->nz);
    for (int l : {p->lx, p->ly, p->lz})
        if (l < 1 || l > FD_MAX_R)
            return fail(PHC_ERR_INVALID, "half width %d outside 1..%d", l, FD_MAX_R);
    const int R = std::max(p->lx, std::max(p->ly, p->lz));
    if (p->xs < 0 || p->xs >= p->nx || p->ys < 0 || p->ys >= p->ny || p->zs < 0 || p->zs >= p->nz)
        return fail(PHC_ERR_INVALID, "source (%d,%d,%d) outside the %d x %d x %d grid", p->xs,
                    p->ys, p->zs, p->nx, p->ny, p->nz);
    if (!p->coefx || !p->coefy || !p->coefz || !p->vp)
        return fail(PHC_ERR_INVALID, "coefficient or vp array is NULL");
    if (p->nsamples < 0 || (p->nsamples > 0 && !p->rhs))
        return fail(PHC_ERR_INVALID, "rhs is NULL with nsamples = %d", p->nsamples);
    if (p->sponge_width < 0 || p->sponge_width > FD_MAX_SW)
        return fail(PHC_ERR_INVALID, "sponge width %d outside 0..%d", p->sponge_width, FD_MAX_SW);
    if (p->arith != PHC_FD_ARITH_REF && p->arith != PHC_FD_ARITH_FP32)
        return fail(PHC_ERR_INVALID, "unknown arithmetic mode %d", p->arith);
    if (p->block_waves != 0 && p->block_waves != 4 && p->block_waves != 8 && p->block_waves != 16)
        return fail(PHC_ERR_INVALID, "block_waves must be 4, 8 or 16 (got %d)", p->block_waves);
    {
        const int ty = p->block_waves ? p->block_waves : ((R == 4) ? 16 : 4);
        const int tj = p->tile_j ? p->tile_j : (p->block_waves ? ty * 2 : ((R == 4) ? 16 : 8));
        if (tj % ty != 0 || !fd_pick_kernel(R, ty, tj / ty, p->arith, true, false))
            return fail(PHC_ERR_INVALID, "no kernel for half width %d, tile_j %d, %d waves per block",
                        R, tj, ty);
    }
    if (p->chunk_i < -4)
        return fail(PHC_ERR_INVALID, "chunk_i %d: planes per chunk, or -1..-4 for the per-XCD schedule", p->chunk_i);
    if (p->k_remainder < -1 || p->k_remainder > 1)
        return fail(PHC_ERR_INVALID, "k_remainder must be -1, 0 or 1 (got %d)", p->k_remainder);
    const double pts = (double)(p->nx + 2 * R) * (p->ny + 2 * R) * (p->nz + 2 * R + 32);
    if (pts > 1.5e10) return fail(PHC_ERR_INVALID, "grid too large (%.3g points)", pts);
    for (int64_t n = 0; n < (int64_t)p->nx * p->ny * p->nz; ++n)
        if (!(p->vp[n] >= 0.f) || !std::isfinite(p->vp[n]))
            return fail(PHC_ERR_INVALID, "vp^2 dt^2 at %lld is negative or not finite",
                        (long long)n);
    return PHC_OK;
}

static phc_status fd_upload_interior(phc_fd *h, const float *host_idx3, float *dev)
{
    // host array in IDX3 order (k fastest, no halo) -> device layout
    std::vector<float> buf((size_t)h->L.total, 0.f);
    for (int i = 0; i < h->nx; ++i)
        for (int j = 0; j < h->ny; ++j)
            std::memcpy(&buf[(size_t)fd_off(h->L, i, j, 0)],
                        host_idx3 + ((int64_t)i * h->ny + j) * h->nz, sizeof(float) * h->nz);
    PHC_HIP(hipMemcpy(dev, buf.data(), sizeof(float) * h->L.total, hipMemcpyHostToDevice));
    return PHC_OK;
}

static float fd_sponge_at(const phc_fd *h, int i, int j, int k)
{
    if (h->sw == 0) return 1.f;
    const int W = h->sw;
    const bool inI = i >= W && i < h->nx - W, inJ = j >= W && j < h->ny - W,
               inK = k >= W && k < h->nz - W;
    float s = 1.f;
    if (inJ && inK && !inI) s = h->hfx[i];
    if (inK && inI && !inJ) s = h->hfy[j];
    if (inJ && inI && !inK) s = h->hfz[k];
    return s;
}

static void fd_fill_args(const phc_fd *h, FdStepArgs &a, int ca, int cb)
{
    std::memset(&a, 0, sizeof(a));
    a.pn = h->lev[cb].as<float>();
    a.pnm1 = h->lev[ca].as<float>();
    a.vp = h->vp.as<float>();
    a.L = h->L;
    std::memcpy(a.cx, h->cx, sizeof(a.cx));
    std::memcpy(a.cy, h->cy, sizeof(a.cy));
    std::memcpy(a.cz, h->cz, sizeof(a.cz));
    a.coef0 = h->coef0;
    a.c0f = (float)h->coef0;
    a.sw = h->sw;
    a.fx = h->fx.as<float>();
    a.fy = h->fy.as<float>();
    a.fz = h->fz.as<float>();
    for (int d = 0; d <= h->sw && d <= FD_MAX_SW; ++d)
        a.stab[d] = (float)std::exp((double)h->alpha * std::pow((double)d, 2.0));
    a.src_off = fd_off(h->L, h->xs, h->ys, h->zs);
    a.inj = 0;
    a.rhs_next = 0.f;
    a.trace = nullptr;
    const int TJ = h->ty * h->nj;
    a.tiles_k = (h->nz + FD_TK - 1) / FD_TK;
    a.tiles_j = (h->ny + TJ - 1) / TJ;
    a.chunk_i = h->chunk_i;
    a.ilo = 0; a.ihi = h->nx;
    a.jlo = 0; a.jhi = h->ny;
    a.klo = 0; a.khi = h->nz;
    if (h->pml_on) {
        // slabs outside the inner box [x3,x4) x [y3,y4) x [z3,z4) (FDTDinit.hpp:176-198)
        a.pml_ex = h->pml_ex.as<float>();
        a.pml_ey = h->pml_ey.as<float>();
        a.pml_ez = h->pml_ez.as<float>();
        a.pml_phi = h->pml_phi.as<float>();
        for (int d = 0; d < 3; ++d) a.pml_hd2[d] = h->pml_hd2[d];
        a.px3 = h->pml_nd[0]; a.px4 = h->nx - h->pml_nd[0];
        a.py3 = h->pml_nd[1]; a.py4 = h->ny - h->pml_nd[1];
        a.pz3 = h->pml_nd[2]; a.pz4 = h->nz - h->pml_nd[2];
    }
}

static int fd_num_blocks(const FdStepArgs &a)
{
    // (with a per-XCD schedule the table's length replaces this count)
    const int chunks = a.chunk_i > 0 ? std::max(1, (a.ihi - a.ilo + a.chunk_i - 1) / a.chunk_i) : 1;
    return a.tiles_k * a.tiles_j * chunks;
}

// Remainder slabs (mode 0: the box's last k tile when it would hold 1..32
// points, and its last j tile when it would hold at most a quarter of its
// TJ rows; 1: whenever either is partial; -1: never): the tiles stop at the
// last full tile and extra blocks update the slabs point by point.
// Returns the number of blocks to launch.
static int fd_plan_blocks(FdStepArgs &a, int mode, int block_threads, int TJ)
{
    a.nb_main = 0;
    a.krem = 0;
    a.jrem = 0;
    const int full = a.khi / FD_TK, rem = a.khi - full * FD_TK;
    const bool usek = mode >= 0 && full > 0 && rem > 0 && FD_TK * full >= a.klo && (mode == 1 || rem <= FD_TK / 2);
    const int rows = a.jhi - a.jlo, fullj = rows / TJ, remj = rows - fullj * TJ;
    const bool usej = mode >= 0 && fullj > 0 && remj > 0 && (mode == 1 || 4 * remj <= TJ);
    if (!usek && !usej) return fd_num_blocks(a);
    if (usek) {
        a.tiles_k = full;
        a.krem = FD_TK * full;
    }
    if (usej) {
        a.tiles_j = fullj;
        a.jrem = a.jlo + TJ * fullj;
    }
    a.nb_main = fd_num_blocks(a);
    const int NI = a.ihi - a.ilo;
    int64_t pts = usek ? (int64_t)rem * rows * NI : 0;
    if (usej) pts += (int64_t)((usek ? a.krem : a.khi) - a.klo) * remj * NI;
    return a.nb_main + (int)((pts + block_threads - 1) / block_threads);
}

static phc_status fd_materialize(phc_fd *h, int level)
{
    if (!h->pending[level]) return PHC_OK;
    if (h->sw > 0) {
        dim3 grid((h->nz + 255) / 256, h->ny, h->nx);
        hipLaunchKernelGGL(fd_damp_kernel, grid, dim3(256), 0, h->stream, h->lev[level].as<float>(),
                           h->L, h->sw, h->fx.as<float>(), h->fy.as<float>(), h->fz.as<float>());
        PHC_HIP(hipGetLastError());
    }
    h->pending[level] = 0;
    return PHC_OK;
}

static void fd_drop_graph(phc_fd *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->g_it0 = h->g_n = h->g_cur = h->g_damp = h->g_nrec = -1;
}

static FdKernelFn fd_handle_kernel(const phc_fd *h)
{
    return fd_pick_kernel(h->R, h->ty, h->nj, h->arith, !h->index64 && h->L.total * 4 < ((int64_t)1 << 32),
                          h->pml_on != 0);
}

// The per-XCD schedule (chunk_i = -m) for the tiles `a` plans: ranges of
// [ilo, ihi) x the tile columns, dealt over as many slots per XCD as the
// kernel keeps resident per CU.  Built on the host and uploaded when its key
// (tiles, planes, split, block shape) changes; never during a capture
// (phc_fd_run prepares it before capturing).
static phc_status fd_prepare_sched(phc_fd *h, const FdStepArgs &a)
{
    const int ncol = a.tiles_k * a.tiles_j;
    const long long key = ((((long long)ncol * 4096 + (a.ihi - a.ilo)) * 8 + (-h->chunk_i)) * 64 + h->ty) * 64 +
                          h->nj + ((long long)a.ilo << 50);
    if (h->sched_key == key && h->sched.ptr) return PHC_OK;
    FdKernelFn fn = fd_handle_kernel(h);
    if (!fn) return fail(PHC_ERR_INVALID, "no kernel for R=%d block_waves=%d rows/thread=%d", h->R, h->ty, h->nj);
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(fn), FD_TK * h->ty, 0) !=
            hipSuccess || per_cu < 1)
        per_cu = 1;
    std::vector<int4> tab;
    build_xcd_schedule(a.ihi - a.ilo, a.ilo, ncol, std::max(h->cus / 8, 1) * per_cu, -h->chunk_i, 1, tab);
    PHC_HIP(hipStreamSynchronize(h->stream));  // queued launches may still read the old table
    fd_drop_graph(h);
    PHC_TRY(h->sched.alloc(sizeof(int4) * tab.size()));
    PHC_HIP(hipMemcpy(h->sched.ptr, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice));
    h->sched_nb = (int)tab.size();
    h->sched_key = key;
    return PHC_OK;
}

// Blocks to launch for `a` (tiles, k-remainder slab, schedule table set in a).
static phc_status fd_plan_launch(phc_fd *h, FdStepArgs &a, int *nb)
{
    *nb = fd_plan_blocks(a, h->krem_eff, FD_TK * h->ty, h->ty * h->nj);
    a.sched = nullptr;
    if (h->chunk_i < 0) {
        const int slab = a.nb_main ? *nb - a.nb_main : 0;
        PHC_TRY(fd_prepare_sched(h, a));
        a.sched = h->sched.as<int4>();
        a.nb_main = slab ? h->sched_nb : 0;
        *nb = h->sched_nb + slab;
    }
    return PHC_OK;
}

static phc_status fd_launch_step(phc_fd *h, FdStepArgs a, bool damp)
{
    FdKernelFn fn = fd_handle_kernel(h);
    if (!fn) return fail(PHC_ERR_INVALID, "no kernel for R=%d block_waves=%d rows/thread=%d", h->R, h->ty, h->nj);
    a.damp_prev = damp ? 1 : 0;
    int nb = 0;
    PHC_TRY(fd_plan_launch(h, a, &nb));
    if (a.rec_n > 0) ++nb;  // the sampling block
    hipLaunchKernelGGL(fn, dim3(nb), dim3(FD_TK, h->ty), 0, h->stream, a);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

// One full step from levels cb (current) and ca (previous, overwritten by the
// new level): one marching launch (with the PML slabs when a PML is set);
// with the learned boundary, its correction pass before and its apply pass
// after the stencil launch.
static phc_status fd_launch_full_step(phc_fd *h, const FdStepArgs &a, bool damp, int ca, int cb,
                                      const float *kslab_in = nullptr, float *kslab_out = nullptr)
{
    if (h->lb_on)
        PHC_HIP(lb_enqueue_correct(h->lbi, h->L, h->lev[cb].as<float>(), h->lev[ca].as<float>(),
                                   h->vp.as<float>(), h->stream, kslab_in, kslab_out));
    PHC_TRY(fd_launch_step(h, a, damp));
    if (h->lb_on) PHC_HIP(lb_enqueue_apply(h->lbi, h->L, h->lev[ca].as<float>(), h->stream));
    return PHC_OK;
}

phc_status phc::fd_field_view(phc_fd *h, FdFieldView *v)
{
    if (!h || !v) return fail(PHC_ERR_INVALID, "NULL argument");
    if (h->lh[0] != LB_R || h->lh[1] != LB_R || h->lh[2] != LB_R)
        return fail(PHC_ERR_INVALID, "the learned boundary needs half width %d on every axis", LB_R);
    for (int l = 0; l <= h->R; ++l)
        if (h->cx[l] != h->cy[l] || h->cx[l] != h->cz[l])
            return fail(PHC_ERR_INVALID, "the learned boundary needs equal coefficients on the three axes");
    if (!(h->cx[1] > 0.f)) return fail(PHC_ERR_INVALID, "c_1 must be positive");
    v->L = h->L;
    v->lev[0] = h->lev[0].as<float>();
    v->lev[1] = h->lev[1].as<float>();
    v->vp = h->vp.as<float>();
    v->sw = h->sw;
    v->c1 = h->cx[1];
    for (int l = 0; l <= LB_R; ++l) v->chat[l] = h->cx[l] / h->cx[1];
    v->coef0 = h->coef0;
    v->stream = h->stream;
    return PHC_OK;
}

extern "C" {

phc_status phc_fd_create(const phc_fd_params *p, phc_fd **out)
{
    if (!out) return fail(PHC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    PHC_TRY(fd_validate(p));
    phc_fd *h = new phc_fd();
    h->nx = p->nx; h->ny = p->ny; h->nz = p->nz;
    h->lh[0] = p->lx; h->lh[1] = p->ly; h->lh[2] = p->lz;
    h->R = std::max(p->lx, std::max(p->ly, p->lz));
    h->xs = p->xs; h->ys = p->ys; h->zs = p->zs;
    h->arith = p->arith;
    h->sw = p->sponge_width;
    h->alpha = p->sponge_alpha;
    std::memset(h->cx, 0, sizeof(h->cx));
    std::memset(h->cy, 0, sizeof(h->cy));
    std::memset(h->cz, 0, sizeof(h->cz));
    // Unequal half widths (FDTDkernels.hpp:16-29 loops to each coef*.size()):
    // the kernels run at the largest one, R, with zero weights past an axis's
    // own width.  Those extra terms add 0 * (finite halo values) = 0 exactly,
    // so every sum keeps the reference's value.
    for (int l = 0; l <= p->lx; ++l) h->cx[l] = p->coefx[l];
    for (int l = 0; l <= p->ly; ++l) h->cy[l] = p->coefy[l];
    for (int l = 0; l <= p->lz; ++l) h->cz[l] = p->coefz[l];
    h->coef0 = p->coef0;
    h->nsamples = p->nsamples;
    // default shape: the fastest measured at 512^3 for radius 4 (16 waves,
    // 16 rows, whole i-extent per block); other radii have one shape
    const int dty = (h->R == 4) ? 16 : 4, dtj = (h->R == 4) ? 16 : 8;
    h->ty = p->block_waves ? p->block_waves : dty;
    h->nj = (p->tile_j ? p->tile_j : (p->block_waves ? 2 * h->ty : dtj)) / h->ty;
    h->chunk_i = p->chunk_i ? p->chunk_i : (h->R == 4 ? 512 : 256);
    h->index64 = p->index64 ? 1 : 0;
    h->krem_mode = h->krem_eff = p->k_remainder;
    h->L = make_fd_layout(h->nx, h->ny, h->nz, h->R);
    phc_status st = PHC_OK;
    auto bail = [&](phc_status s) { phc_fd_destroy(h); return s; };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PHC_ERR_HIP, "hipStreamCreate failed"));
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            h->cus = cus;
    }
    const size_t bytes = sizeof(float) * (size_t)h->L.total;
    if ((st = h->lev[0].alloc(bytes)) || (st = h->lev[1].alloc(bytes)) || (st = h->vp.alloc(bytes)))
        return bail(st);
    // both levels (and their halo and padding) start at init_value
    {
        std::vector<float> init((size_t)h->L.total, p->init_value);
        if (hipMemcpy(h->lev[0].ptr, init.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->lev[1].ptr, init.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(PHC_ERR_HIP, "initial upload failed"));
    }
    if ((st = fd_upload_interior(h, p->vp, h->vp.as<float>()))) return bail(st);
    h->hvp.assign(p->vp, p->vp + (size_t)h->nx * h->ny * h->nz);
    fd_axis_damping(h->nx, h->sw, h->alpha, h->hfx);
    fd_axis_damping(h->ny, h->sw, h->alpha, h->hfy);
    fd_axis_damping(h->nz, h->sw, h->alpha, h->hfz);
    if ((st = h->fx.alloc(sizeof(float) * h->nx)) || (st = h->fy.alloc(sizeof(float) * h->ny)) ||
        (st = h->fz.alloc(sizeof(float) * h->nz)))
        return bail(st);
    if (hipMemcpy(h->fx.ptr, h->hfx.data(), sizeof(float) * h->nx, hipMemcpyHostToDevice) ||
        hipMemcpy(h->fy.ptr, h->hfy.data(), sizeof(float) * h->ny, hipMemcpyHostToDevice) ||
        hipMemcpy(h->fz.ptr, h->hfz.data(), sizeof(float) * h->nz, hipMemcpyHostToDevice))
        return bail(fail(PHC_ERR_HIP, "sponge upload failed"));
    h->hrhs.assign(p->rhs, p->rhs + h->nsamples);
    if (h->nsamples > 0) {
        if ((st = h->rhs.alloc(sizeof(float) * h->nsamples))) return bail(st);
        if (hipMemcpy(h->rhs.ptr, p->rhs, sizeof(float) * h->nsamples, hipMemcpyHostToDevice))
            return bail(fail(PHC_ERR_HIP, "rhs upload failed"));
    }
    *out = h;
    return PHC_OK;
}

phc_status phc_fd_destroy(phc_fd *h)
{
    if (!h) return PHC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    fd_drop_graph(h);
    for (DevBuf *b : {&h->lev[0], &h->lev[1], &h->vp, &h->fx, &h->fy, &h->fz, &h->rhs, &h->trace,
                      &h->rec, &h->recoff, &h->lb_w, &h->lb_corr, &h->lb_kvcs, &h->lb_kslab, &h->pml_ex,
                      &h->pml_ey, &h->pml_ez, &h->pml_phi})
        b->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PHC_OK;
}

phc_status phc_fd_add_rhs(phc_fd *h, int it, int cb)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (cb < 0 || cb > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", cb);
    if (it < 0 || it >= h->nsamples)
        return fail(PHC_ERR_INVALID, "time sample %d outside 0..%d", it, h->nsamples - 1);
    PHC_TRY(fd_materialize(h, cb));
    hipLaunchKernelGGL(fd_add_rhs_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[cb].as<float>(),
                       h->vp.as<float>(), fd_off(h->L, h->xs, h->ys, h->zs), h->rhs.as<float>(), it);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

phc_status phc_fd_add_rhs_value(phc_fd *h, int cb, int i, int j, int k, float rhs)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (cb < 0 || cb > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", cb);
    if (i < 0 || i >= h->nx || j < 0 || j >= h->ny || k < 0 || k >= h->nz)
        return fail(PHC_ERR_INVALID, "source (%d,%d,%d) outside the %d x %d x %d grid", i, j, k, h->nx, h->ny,
                    h->nz);
    if (!std::isfinite(rhs)) return fail(PHC_ERR_INVALID, "source value is not finite");
    PHC_TRY(fd_materialize(h, cb));
    hipLaunchKernelGGL(fd_add_value_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[cb].as<float>(),
                       h->vp.as<float>(), fd_off(h->L, i, j, k), rhs);
    PHC_HIP(hipGetLastError());
    return PHC_OK;
}

phc_status phc_fd_compute_one_step_pml(phc_fd *h, int ca, int cb)
{
    RoctxRange rr_("phc_fd_compute_one_step_pml");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (!h->pml_on) return fail(PHC_ERR_STATE, "no PML set (phc_fd_set_pml)");
    if (ca < 0 || ca > 1 || cb < 0 || cb > 1 || ca == cb)
        return fail(PHC_ERR_INVALID, "levels ca=%d cb=%d must be 0 and 1", ca, cb);
    FdStepArgs a;
    fd_fill_args(h, a, ca, cb);
    PHC_TRY(fd_launch_full_step(h, a, false, ca, cb));
    return PHC_OK;
}

phc_status phc_fd_compute_one_step_sb(phc_fd *h, int ca, int cb)
{
    RoctxRange rr_("phc_fd_compute_one_step_sb");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (h->pml_on) return fail(PHC_ERR_STATE, "a PML is set: use phc_fd_compute_one_step_pml");
    if (ca < 0 || ca > 1 || cb < 0 || cb > 1 || ca == cb)
        return fail(PHC_ERR_INVALID, "levels ca=%d cb=%d must be 0 and 1", ca, cb);
    PHC_TRY(fd_materialize(h, cb));
    FdStepArgs a;
    fd_fill_args(h, a, ca, cb);
    PHC_TRY(fd_launch_full_step(h, a, h->pending[ca] != 0, ca, cb));
    h->pending[ca] = 0;
    h->pending[cb] = (h->sw > 0) ? 1 : 0;
    return PHC_OK;
}

phc_status phc_fd_set_receivers(phc_fd *h, int nrec, const int *ijk)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (nrec < 0 || (nrec > 0 && !ijk)) return fail(PHC_ERR_INVALID, "bad receiver list");
    std::vector<int64_t> offs(nrec);
    for (int r = 0; r < nrec; ++r) {
        const int i = ijk[3 * r], j = ijk[3 * r + 1], k = ijk[3 * r + 2];
        if (i < 0 || i >= h->nx || j < 0 || j >= h->ny || k < 0 || k >= h->nz)
            return fail(PHC_ERR_INVALID, "receiver %d at (%d,%d,%d) is outside the grid", r, i, j, k);
        offs[r] = fd_off(h->L, i, j, k);
    }
    PHC_HIP(hipStreamSynchronize(h->stream));
    fd_drop_graph(h);
    h->rec_offs = offs;
    h->nrec = nrec;
    if (nrec > 0) {
        PHC_TRY(h->recoff.alloc(sizeof(int64_t) * nrec));
        PHC_HIP(hipMemcpy(h->recoff.ptr, offs.data(), sizeof(int64_t) * nrec, hipMemcpyHostToDevice));
    }
    return PHC_OK;
}


} // extern "C"

// Enqueue the whole run on h->stream: prologue injection of rhs[it0] into the
// current level, then one fused step per time sample (injection of the next
// sample folded into the store), plus receiver sampling.  With `ev`, events
// bracket every stencil launch (timing mode).
static phc_status fd_enqueue_run(phc_fd *h, int it0, int nsteps, int c0, bool damp0, hipEvent_t *ev)
{
    hipLaunchKernelGGL(fd_add_rhs_kernel, dim3(1), dim3(64), 0, h->stream, h->lev[c0].as<float>(),
                       h->vp.as<float>(), fd_off(h->L, h->xs, h->ys, h->zs), h->rhs.as<float>(),
                       it0);
    PHC_HIP(hipGetLastError());
    int c = c0;
    for (int s = 0; s < nsteps; ++s) {
        const int it = it0 + s;
        const int ca = 1 - c, cb = c;
        FdStepArgs a;
        fd_fill_args(h, a, ca, cb);
        a.trace = h->trace.as<float>() + s;
        if (s + 1 < nsteps) {
            a.inj = 1;
            a.rhs_next = h->hrhs[it + 1];
        }
        if (h->nrec > 0 && s > 0) {  // the previous step's receivers, by this launch
            a.rec_off = h->recoff.as<int64_t>();
            a.rec_out = h->rec.as<float>();
            a.rec_src = h->trace.as<float>();
            a.rec_n = h->nrec;
            a.rec_stride = nsteps;
            a.rec_step = s - 1;
        }
        if (ev) PHC_HIP(hipEventRecord(ev[2 * s], h->stream));
        // learned layer: step s's correction pass hands its k-face planes to
        // step s + 1's through store s % 2 (the first step reads the grid)
        float *kout = h->lb_on ? h->lbi.kslab[s & 1] : nullptr;
        const float *kin = h->lb_on && s > 0 ? h->lbi.kslab[(s + 1) & 1] : nullptr;
        PHC_TRY(fd_launch_full_step(h, a, s == 0 ? damp0 : (h->sw > 0), ca, cb, kin, kout));
        if (ev) PHC_HIP(hipEventRecord(ev[2 * s + 1], h->stream));
        c = ca;
    }
    if (h->nrec > 0) {  // the last step's receivers
        hipLaunchKernelGGL(fd_sample_kernel, dim3((h->nrec + 255) / 256), dim3(256), 0, h->stream,
                           h->lev[c].as<float>(), h->recoff.as<int64_t>(), h->nrec, h->rec.as<float>(), nsteps,
                           nsteps - 1, fd_off(h->L, h->xs, h->ys, h->zs), h->trace.as<float>());
        PHC_HIP(hipGetLastError());
    }
    return PHC_OK;
}

extern "C" {

phc_status phc_fd_run(phc_fd *h, int it0, int nsteps, int *cur)
{
    RoctxRange rr_("phc_fd_run");
    if (!h || !cur) return fail(PHC_ERR_INVALID, "handle or cur is NULL");
    if (*cur < 0 || *cur > 1) return fail(PHC_ERR_INVALID, "cur level %d is not 0 or 1", *cur);
    if (nsteps < 0) return fail(PHC_ERR_INVALID, "negative step count");
    if (nsteps == 0) return PHC_OK;
    if (it0 < 0 || it0 + nsteps > h->nsamples)
        return fail(PHC_ERR_INVALID, "steps %d..%d outside the %d source samples", it0,
                    it0 + nsteps - 1, h->nsamples);
    const int c0 = *cur;
    PHC_TRY(fd_materialize(h, c0));
    const bool damp0 = h->pending[1 - c0] != 0;
    if (h->trace.bytes < sizeof(float) * nsteps) {
        PHC_HIP(hipStreamSynchronize(h->stream));
        fd_drop_graph(h);
        PHC_TRY(h->trace.alloc(sizeof(float) * nsteps));
    }
    if (h->nrec > 0 && h->rec.bytes < sizeof(float) * (size_t)nsteps * h->nrec) {
        PHC_HIP(hipStreamSynchronize(h->stream));
        fd_drop_graph(h);
        PHC_TRY(h->rec.alloc(sizeof(float) * (size_t)nsteps * h->nrec));
    }
    if (h->timing) {
        std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
        for (auto &e : ev) PHC_HIP(hipEventCreate(&e));
        phc_status st = fd_enqueue_run(h, it0, nsteps, c0, damp0, ev.data());
        if (st == PHC_OK) {
            PHC_HIP(hipStreamSynchronize(h->stream));
            double tot = 0.0;
            for (int s = 0; s < nsteps; ++s) {
                float ms = 0.f;
                PHC_HIP(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
                tot += ms;
            }
            h->last_ms = tot / nsteps;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        PHC_TRY(st);
    } else {
        const int damp_key = damp0 ? 1 : 0;
        if (!(h->gexec && h->g_it0 == it0 && h->g_n == nsteps && h->g_cur == c0 &&
              h->g_damp == damp_key && h->g_nrec == h->nrec)) {
            fd_drop_graph(h);
            {   // the schedule table is uploaded now, not inside the capture
                FdStepArgs a;
                fd_fill_args(h, a, 1 - c0, c0);
                int nb = 0;
                PHC_TRY(fd_plan_launch(h, a, &nb));
            }
            hipGraph_t graph = nullptr;
            PHC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            phc_status st = fd_enqueue_run(h, it0, nsteps, c0, damp0, nullptr);
            hipError_t ce = hipStreamEndCapture(h->stream, &graph);
            if (st != PHC_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (ce != hipSuccess) return fail(PHC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                h->gexec = nullptr;
                return fail(PHC_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
            }
            h->g_it0 = it0; h->g_n = nsteps; h->g_cur = c0; h->g_damp = damp_key; h->g_nrec = h->nrec;
        }
        PHC_HIP(hipGraphLaunch(h->gexec, h->stream));
    }
    const int cend = (nsteps % 2 == 0) ? c0 : 1 - c0;
    h->pending[cend] = 0;
    h->pending[1 - cend] = (h->sw > 0) ? 1 : 0;
    h->last_n = nsteps;
    h->last_nrec = h->nrec;
    *cur = cend;
    return PHC_OK;
}

phc_status phc_fd_get_traces(phc_fd *h, float *src, float *rec)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    if (src && h->last_n > 0)
        PHC_HIP(hipMemcpy(src, h->trace.ptr, sizeof(float) * h->last_n, hipMemcpyDeviceToHost));
    if (rec && h->last_n > 0 && h->last_nrec > 0)
        PHC_HIP(hipMemcpy(rec, h->rec.ptr, sizeof(float) * (size_t)h->last_n * h->last_nrec,
                          hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_fd_get_field(phc_fd *h, int level, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "handle or host buffer is NULL");
    if (level < 0 || level > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", level);
    PHC_HIP(hipStreamSynchronize(h->stream));
    std::vector<float> buf((size_t)h->L.total);
    PHC_HIP(hipMemcpy(buf.data(), h->lev[level].ptr, sizeof(float) * h->L.total, hipMemcpyDeviceToHost));
    // IDX3_l order with the reference's per-axis halos (FDTDmacros.hpp:14)
    const int LX = h->lh[0], LY = h->lh[1], LZ = h->lh[2];
    const int NYe = h->ny + 2 * LY, NZe = h->nz + 2 * LZ;
    for (int i = -LX; i < h->nx + LX; ++i)
        for (int j = -LY; j < h->ny + LY; ++j) {
            float *dst = host + ((int64_t)(i + LX) * NYe + (j + LY)) * NZe;
            const float *srcp = &buf[(size_t)fd_off(h->L, i, j, -LZ)];
            std::memcpy(dst, srcp, sizeof(float) * NZe);
            const bool row_in = i >= 0 && i < h->nx && j >= 0 && j < h->ny;
            if (h->pending[level] && row_in)
                for (int k = 0; k < h->nz; ++k) dst[k + LZ] = dst[k + LZ] * fd_sponge_at(h, i, j, k);
        }
    return PHC_OK;
}

phc_status phc_fd_set_field(phc_fd *h, int level, const float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "handle or host buffer is NULL");
    if (level < 0 || level > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", level);
    PHC_HIP(hipStreamSynchronize(h->stream));
    std::vector<float> buf((size_t)h->L.total);
    PHC_HIP(hipMemcpy(buf.data(), h->lev[level].ptr, sizeof(float) * h->L.total, hipMemcpyDeviceToHost));
    const int LX = h->lh[0], LY = h->lh[1], LZ = h->lh[2];
    const int NYe = h->ny + 2 * LY, NZe = h->nz + 2 * LZ;
    for (int i = -LX; i < h->nx + LX; ++i)
        for (int j = -LY; j < h->ny + LY; ++j)
            std::memcpy(&buf[(size_t)fd_off(h->L, i, j, -LZ)],
                        host + ((int64_t)(i + LX) * NYe + (j + LY)) * NZe, sizeof(float) * NZe);
    PHC_HIP(hipMemcpy(h->lev[level].ptr, buf.data(), sizeof(float) * h->L.total, hipMemcpyHostToDevice));
    h->pending[level] = 0;
    return PHC_OK;
}

phc_status phc_fd_read_point(phc_fd *h, int level, int i, int j, int k, float *v)
{
    if (!h || !v) return fail(PHC_ERR_INVALID, "handle or output is NULL");
    if (level < 0 || level > 1) return fail(PHC_ERR_INVALID, "level %d is not 0 or 1", level);
    const int LX = h->lh[0], LY = h->lh[1], LZ = h->lh[2];
    if (i < -LX || i >= h->nx + LX || j < -LY || j >= h->ny + LY || k < -LZ || k >= h->nz + LZ)
        return fail(PHC_ERR_INVALID, "point (%d,%d,%d) outside the grid and halo", i, j, k);
    PHC_HIP(hipStreamSynchronize(h->stream));
    float x = 0.f;
    PHC_HIP(hipMemcpy(&x, h->lev[level].as<float>() + fd_off(h->L, i, j, k), sizeof(float),
                      hipMemcpyDeviceToHost));
    const bool in = i >= 0 && i < h->nx && j >= 0 && j < h->ny && k >= 0 && k < h->nz;
    if (h->pending[level] && in) x = x * fd_sponge_at(h, i, j, k);
    *v = x;
    return PHC_OK;
}

phc_status phc_fd_get_slice_xz(phc_fd *h, int level, int j, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "handle or host buffer is NULL");
    if (j < 0 || j >= h->ny) return fail(PHC_ERR_INVALID, "slice j=%d outside 0..%d", j, h->ny - 1);
    const int LX = h->lh[0], LY = h->lh[1], LZ = h->lh[2];
    std::vector<float> full((size_t)(h->nx + 2 * LX) * (h->ny + 2 * LY) * (h->nz + 2 * LZ));
    PHC_TRY(phc_fd_get_field(h, level, full.data()));
    const int NYe = h->ny + 2 * LY, NZe = h->nz + 2 * LZ;
    for (int k = 0; k < h->nz; ++k)
        for (int i = 0; i < h->nx; ++i)
            host[(int64_t)k * h->nx + i] = full[((int64_t)(i + LX) * NYe + (j + LY)) * NZe + (k + LZ)];
    return PHC_OK;
}

phc_status phc_fd_check_finite(phc_fd *h, int level)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    std::vector<float> full((size_t)(h->nx + 2 * h->lh[0]) * (h->ny + 2 * h->lh[1]) * (h->nz + 2 * h->lh[2]));
    PHC_TRY(phc_fd_get_field(h, level, full.data()));
    for (size_t n = 0; n < full.size(); ++n)
        if (!std::isfinite(full[n]))
            return fail(PHC_ERR_NONFINITE, "level %d holds a non-finite value at %zu", level, n);
    return PHC_OK;
}

phc_status phc_fd_last_run_kernel_ms(phc_fd *h, double *ms)
{
    if (!h || !ms) return fail(PHC_ERR_INVALID, "handle or output is NULL");
    *ms = h->last_ms;
    return PHC_OK;
}

phc_status phc_fd_set_timing(phc_fd *h, int enable)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    h->timing = enable ? 1 : 0;
    return PHC_OK;
}

phc_status phc_fd_set_pml(phc_fd *h, const int *ndamp, const float *etax, const float *etay, const float *etaz,
                          const float *hd2)
{
    if (!h || !ndamp) return fail(PHC_ERR_INVALID, "handle or ndamp is NULL");
    const int n[3] = {h->nx, h->ny, h->nz};
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        if (ndamp[a] < 0 || 2 * ndamp[a] > n[a])
            return fail(PHC_ERR_INVALID, "PML width %d on axis %d does not fit %d points", ndamp[a], a, n[a]);
        any = any || ndamp[a] > 0;
    }
    PHC_HIP(hipStreamSynchronize(h->stream));
    fd_drop_graph(h);
    if (!any) {
        h->pml_on = 0;
        h->pml_nd[0] = h->pml_nd[1] = h->pml_nd[2] = 0;
        for (DevBuf *b : {&h->pml_ex, &h->pml_ey, &h->pml_ez, &h->pml_phi}) b->release();
        return PHC_OK;
    }
    if (!etax || !etay || !etaz || !hd2) return fail(PHC_ERR_INVALID, "eta profile or hd2 is NULL");
    if (h->sw != 0) return fail(PHC_ERR_INVALID, "the PML replaces the sponge: sponge_width must be 0");
    if (h->lb_on) return fail(PHC_ERR_STATE, "a learned boundary is attached");
    const float *prof[3] = {etax, etay, etaz};
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < n[a] + 2; ++i)
            if (!std::isfinite(prof[a][i])) return fail(PHC_ERR_INVALID, "eta profile %d is not finite", a);
    DevBuf *dst[3] = {&h->pml_ex, &h->pml_ey, &h->pml_ez};
    for (int a = 0; a < 3; ++a) {
        PHC_TRY(dst[a]->alloc(sizeof(float) * (n[a] + 2)));
        PHC_HIP(hipMemcpy(dst[a]->ptr, prof[a], sizeof(float) * (n[a] + 2), hipMemcpyHostToDevice));
        h->pml_nd[a] = ndamp[a];
        h->pml_hd2[a] = hd2[a];
    }
    PHC_TRY(h->pml_phi.alloc(sizeof(float) * (size_t)h->L.total));
    PHC_HIP(hipMemset(h->pml_phi.ptr, 0, sizeof(float) * (size_t)h->L.total));
    h->pml_on = 1;
    return PHC_OK;
}

phc_status phc_fd_set_learned_boundary(phc_fd *h, const phc_lb *lb, float halo, float layer_damping)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    fd_drop_graph(h);
    if (!lb) {
        h->lb_on = 0;
        h->lb_w.release();
        h->lb_corr.release();
        h->lb_kvcs.release();
        h->lb_kslab.release();
        return PHC_OK;
    }
    FdFieldView fv;
    PHC_TRY(fd_field_view(h, &fv));
    if (h->sw != 0) return fail(PHC_ERR_INVALID, "the learned boundary replaces the sponge: sponge_width must be 0");
    if (h->pml_on) return fail(PHC_ERR_STATE, "a PML is set");
    const int n[3] = {h->nx, h->ny, h->nz}, src[3] = {h->xs, h->ys, h->zs};
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 4 * LB_W || n[a] > 4000)
            return fail(PHC_ERR_INVALID, "grid side %d outside %d..4000", n[a], 4 * LB_W);
        if (src[a] < LB_W || src[a] >= n[a] - LB_W)
            return fail(PHC_ERR_INVALID, "the source must lie inside the %d-point learned layer", LB_W);
    }
    if (!std::isfinite(halo)) return fail(PHC_ERR_INVALID, "halo value is not finite");
    if (!(layer_damping <= 0.f) || layer_damping < -1.f)
        return fail(PHC_ERR_INVALID, "layer damping %g outside [-1, 0]", layer_damping);
    const int64_t ncorr = lb_corr_layout(h->L, h->lbi.corr_off);
    PHC_TRY(h->lb_w.alloc(sizeof(uint16_t) * LB_FACES * LB_NP));
    PHC_TRY(h->lb_corr.alloc(sizeof(float) * (size_t)ncorr));
    PHC_TRY(lb_export_bf16(lb, h->lb_w.as<uint16_t>(), h->stream));
    h->lbi.w = h->lb_w.as<uint16_t>();
    h->lbi.corr = h->lb_corr.as<float>();
    h->lbi.halo = halo;
    h->lbi.c1 = fv.c1;
    for (int l = 0; l <= LB_R; ++l) h->lbi.chat[l] = fv.chat[l];
    for (int d = 0; d < LB_W; ++d)
        h->lbi.ldamp[d] = (float)std::exp((double)layer_damping * (double)(LB_W - d) * (LB_W - d));
    // the k-faces' compact velocity terms (vp is fixed after create) and the
    // two stores the fused run's correction passes hand their planes on in
    const int64_t npk = (int64_t)h->nx * h->ny;
    PHC_TRY(h->lb_kvcs.alloc(sizeof(float) * (size_t)(2 * 2 * LB_W * npk)));
    PHC_TRY(h->lb_kslab.alloc(sizeof(float) * (size_t)(2 * 2 * (LB_W + LB_R) * npk)));
    h->lbi.npts_k = (int)npk;
    h->lbi.kslab[0] = h->lb_kslab.as<float>();
    h->lbi.kslab[1] = h->lbi.kslab[0] + 2 * (LB_W + LB_R) * npk;
    h->lbi.kvcs = nullptr;
    PHC_HIP(lb_enqueue_kface_vp(h->lbi, h->L, h->vp.as<float>(), h->lb_kvcs.as<float>(), h->stream));
    h->lbi.kvcs = h->lb_kvcs.as<float>();
    h->lb_on = 1;
    return PHC_OK;
}

phc_status phc_fd_lb_corrections(phc_fd *h, int cur, float *host)
{
    if (!h || !host) return fail(PHC_ERR_INVALID, "handle or host buffer is NULL");
    if (!h->lb_on) return fail(PHC_ERR_STATE, "no learned boundary attached");
    if (cur < 0 || cur > 1) return fail(PHC_ERR_INVALID, "cur level %d is not 0 or 1", cur);
    PHC_HIP(lb_enqueue_correct(h->lbi, h->L, h->lev[cur].as<float>(), h->lev[1 - cur].as<float>(),
                               h->vp.as<float>(), h->stream));
    PHC_HIP(hipStreamSynchronize(h->stream));
    PHC_HIP(hipMemcpy(host, h->lb_corr.ptr, sizeof(float) * h->lbi.corr_off[LB_FACES], hipMemcpyDeviceToHost));
    return PHC_OK;
}

phc_status phc_fd_synchronize(phc_fd *h)
{
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    PHC_HIP(hipStreamSynchronize(h->stream));
    return PHC_OK;
}

} // extern "C"

// ================================================================ autotuner
// Times every block shape / chunk length the library has a kernel for on the
// handle's grid (on scratch copies of the two levels, so the simulation state
// is untouched) and keeps the fastest; with k_remainder 0 it also times the
// last partial k tile as a tile (-1) and keeps the faster of the two.  Results
// are cached in a text file, one line per (grid, radius, arithmetic, step
// variant, requested k mode):
// "fd nx ny nz R arith pml kmode waves tile_j chunk_i kmode_used ms".
namespace phc {

static bool fd_tune_lookup(const char *path, const phc_fd *h, int *ty, int *tj, int *ci, int *km_used)
{
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    bool hit = false;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {  // lines with other tags are skipped whole
        char tag[8];
        int nx, ny, nz, R, ar, pml, km, a, b, c, ku;
        double ms;
        if (std::sscanf(line, "%7s %d %d %d %d %d %d %d %d %d %d %d %lf", tag, &nx, &ny, &nz, &R, &ar, &pml, &km, &a,
                        &b, &c, &ku, &ms) != 13)
            continue;
        if (!std::strcmp(tag, "fd") && nx == h->nx && ny == h->ny && nz == h->nz && R == h->R && ar == h->arith &&
            pml == h->pml_on && km == h->krem_mode && a > 0 && b % a == 0 && c >= -4 && c != 0 && ku >= -1 &&
            ku <= 1 &&
            (ku == km || km == 0) && fd_pick_kernel(h->R, a, b / a, h->arith, true, false)) {
            *ty = a; *tj = b; *ci = c; *km_used = ku;
            hit = true;
        }
    }
    std::fclose(f);
    return hit;
}

} // namespace phc

extern "C" phc_status phc_fd_autotune(phc_fd *h, int steps_per_trial, const char *cache_path,
                                      int *tile_j, int *block_waves, int *chunk_i)
{
    RoctxRange rr_("phc_fd_autotune");
    if (!h) return fail(PHC_ERR_INVALID, "handle is NULL");
    if (steps_per_trial < 1) return fail(PHC_ERR_INVALID, "steps_per_trial must be >= 1");
    int ty = 0, tj = 0, ci = 0, ku = 0;
    if (fd_tune_lookup(cache_path, h, &ty, &tj, &ci, &ku)) {
        h->ty = ty; h->nj = tj / ty; h->chunk_i = ci; h->krem_eff = ku;
    } else {
        const size_t bytes = sizeof(float) * (size_t)h->L.total;
        DevBuf s0, s1;
        PHC_TRY(s0.alloc(bytes));
        PHC_TRY(s1.alloc(bytes));
        PHC_HIP(hipMemcpyAsync(s0.ptr, h->lev[0].ptr, bytes, hipMemcpyDeviceToDevice, h->stream));
        PHC_HIP(hipMemcpyAsync(s1.ptr, h->lev[1].ptr, bytes, hipMemcpyDeviceToDevice, h->stream));
        DevBuf s2;  // the PML variant writes phi: its trials run on a copy
        if (h->pml_on) {
            PHC_TRY(s2.alloc(bytes));
            PHC_HIP(hipMemcpyAsync(s2.ptr, h->pml_phi.ptr, bytes, hipMemcpyDeviceToDevice, h->stream));
        }
        hipEvent_t e0, e1;
        PHC_HIP(hipEventCreate(&e0));
        PHC_HIP(hipEventCreate(&e1));
        const int shapes[][2] = {{4, 8}, {4, 16}, {8, 8}, {8, 16}, {16, 16}, {16, 32}};
        const int chunks[] = {64, 128, 256, 512, 1024, -1, -2, -3, -4};
        double best = 1e30;
        int bty = h->ty, bnj = h->nj, bci = h->chunk_i, bku = h->krem_eff;
        phc_status st = PHC_OK;
        // with mode 0 the plain partial tiles (-1) are timed as well
        const bool both = h->krem_mode == 0;
        const int modes[2] = {h->krem_mode, -1};
        for (int mi = 0; mi < (both ? 2 : 1) && st == PHC_OK; ++mi)
        for (auto &sh : shapes) {
            const int ku = modes[mi];
            if (!fd_pick_kernel(h->R, sh[0], sh[1] / sh[0], h->arith, true, false)) continue;
            for (int c : chunks) {
                if (c > h->nx && c != chunks[0]) continue;
                h->ty = sh[0]; h->nj = sh[1] / sh[0]; h->chunk_i = c < 0 ? c : std::min(c, h->nx); h->krem_eff = ku;
                float ms = 0.f;
                for (int rep = 0; rep < 2 && st == PHC_OK; ++rep) {  // first pass warms up
                    if (hipEventRecord(e0, h->stream) != hipSuccess) st = fail(PHC_ERR_HIP, "event record failed");
                    for (int s = 0; s < steps_per_trial && st == PHC_OK; ++s) {
                        FdStepArgs a;
                        fd_fill_args(h, a, 0, 1);
                        a.pn = s1.as<float>();
                        a.pnm1 = s0.as<float>();
                        if (s & 1) { a.pn = s0.as<float>(); a.pnm1 = s1.as<float>(); }
                        if (h->pml_on) a.pml_phi = s2.as<float>();
                        st = fd_launch_step(h, a, false);
                    }
                    if (st == PHC_OK && (hipEventRecord(e1, h->stream) != hipSuccess ||
                                         hipEventSynchronize(e1) != hipSuccess ||
                                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        st = fail(PHC_ERR_HIP, "trial timing failed");
                }
                if (st != PHC_OK) break;
                if (ms < best) { best = ms; bty = h->ty; bnj = h->nj; bci = h->chunk_i; bku = h->krem_eff; }
            }
            if (st != PHC_OK) break;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->ty = bty; h->nj = bnj; h->chunk_i = bci; h->krem_eff = bku;
        fd_drop_graph(h);
        PHC_TRY(st);
        if (cache_path) {
            FILE *f = std::fopen(cache_path, "a");
            if (f) {
                std::fprintf(f, "fd %d %d %d %d %d %d %d %d %d %d %d %.5f\n", h->nx, h->ny, h->nz, h->R,
                             h->arith, h->pml_on, h->krem_mode, bty, bty * bnj, bci, bku, best / steps_per_trial);
                std::fclose(f);
            }
        }
    }
    fd_drop_graph(h);
    if (tile_j) *tile_j = h->ty * h->nj;
    if (block_waves) *block_waves = h->ty;
    if (chunk_i) *chunk_i = h->chunk_i;
    return PHC_OK;
}
