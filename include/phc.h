/*
 * phc.h -- C ABI of libphc.so, the MI355X (gfx950) time-stepping hot path of
 * the numpex/proxy-geos-hc acoustic proxies.
 *
 * Plain pointers and sizes only.  Every entry point returns a phc_status
 * (0 = ok, < 0 = error) and never throws; phc_last_error() returns a
 * thread-local message for the last failing call on that thread.  A handle
 * owns all of its device memory and one HIP stream; host arrays passed in are
 * copied at create time and stay owned by the caller.  One host thread per
 * handle; different handles are independent.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree).  The reference has no C ABI: these are the calls its
 * C++ classes (FDTDKernel, FDTDUtils, SEMsolver) would make across an FFI; the
 * C++ drop-in classes in proxy-geos-hc_amd/host/ are written on top of them.
 */
#ifndef PHC_H
#define PHC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int phc_status;
#define PHC_OK 0
#define PHC_ERR_INVALID (-1)   /* bad argument or unsupported configuration */
#define PHC_ERR_HIP (-2)       /* a HIP runtime call failed */
#define PHC_ERR_NOMEM (-3)     /* device allocation failed */
#define PHC_ERR_NONFINITE (-4) /* a NaN/Inf was detected by phc_*_check_finite */
#define PHC_ERR_STATE (-5)     /* call not valid in the handle's current state */

/* Thread-local text of the last error (never NULL). */
const char *phc_last_error(void);
/* Library version string. */
const char *phc_version(void);
/* Number of visible HIP devices (0 if none); never fails. */
int phc_device_count(void);
/* hipDeviceSynchronize on the current device. */
phc_status phc_device_synchronize(void);
/* Plain streaming copy of `bytes` (device to device), iters times; mean ms per
 * copy.  Calibrates the HBM byte counters and the practical copy bandwidth.
 * No reference counterpart (measurement utility). */
phc_status phc_copy_bandwidth(size_t bytes, int iters, double *ms_per_copy);

/* ====================================================================== FD
 * Replaces FDTDKernel (src/fd/utilsFDTD/FDTDkernels.hpp:4-131) and the
 * device-side part of FDTDInit::init_models/defineSpongeBoundary
 * (src/fd/utilsFDTD/FDTDinit.hpp:249-396) and FDTDUtils::output
 * (src/fd/utilsFDTD/FDTDutils.hpp:193-232).
 */
typedef struct phc_fd phc_fd;

/* Arithmetic of one point update (FDTDkernels.hpp:15-31). */
#define PHC_FD_ARITH_REF 0  /* reference rounding: float lap sums, FP64 combine */
#define PHC_FD_ARITH_FP32 1 /* all-FP32 with fused multiply-add */

typedef struct {
    /* FDTDGRIDS (src/fd/utilsFDTD/FDTDdata.hpp:6-23) */
    int nx, ny, nz;   /* interior points per axis */
    int lx, ly, lz;   /* stencil half widths (halo widths); must be equal, 1..6 */
    int xs, ys, zs;   /* source point (interior coordinates) */
    /* FDTDMODELS (FDTDdata.hpp:25-40) */
    const float *coefx, *coefy, *coefz; /* [l+1] each, coef[0] unused by the step */
    double coef0;                       /* centre coefficient (FDTDinit.hpp:227) */
    const float *vp;   /* vp^2 dt^2 per interior point, IDX3 order (k fastest) */
    const float *rhs;  /* RHSTerm[nsamples] */
    int nsamples;
    /* sponge (FDTDinit.hpp:319-396): face slabs of width sponge_width with
     * exp(alpha (L-i)^2); edges and corners 1.  sponge_width = 0: no sponge. */
    int sponge_width;
    float sponge_alpha;
    float init_value;  /* initial value of both levels and of the halo (1e-6) */
    int arith;         /* PHC_FD_ARITH_* */
    /* kernel tiling; 0 selects the built-in default for the radius */
    int tile_j;        /* output rows per block along j (block_waves x 1, 2 or 4) */
    int block_waves;   /* wavefronts per block: 4, 8 or 16 */
    int chunk_i;       /* planes per block along the marching axis i */
    int index64;       /* 1: force 64-bit element offsets (default: 32-bit when
                        * every array is below 4 GiB) */
} phc_fd_params;

phc_status phc_fd_create(const phc_fd_params *p, phc_fd **out);
phc_status phc_fd_destroy(phc_fd *h);

/* FDTDKernel::addRHS (FDTDkernels.hpp:121-128): level[cb][src] += vp[src]*rhs[it]. */
phc_status phc_fd_add_rhs(phc_fd *h, int it, int cb);
/* FDTDKernel::computeOneStepSB (FDTDkernels.hpp:75-84) = inner3D + applySponge:
 * level ca <- sponge * (2 level[cb] - level[ca] + vp (coef0 p + lap p)),
 * level cb <- sponge * level cb.  The damping of cb is deferred to the next
 * read of that level (same values, one pass fewer). */
phc_status phc_fd_compute_one_step_sb(phc_fd *h, int ca, int cb);

/* The FD time loop of FDTDmain.cpp:55-81 for steps it0 .. it0+nsteps-1:
 * addRHS(cur) -> computeOneStepSB(prev, cur) -> swap, with the injection fused
 * into the preceding step's store and the whole run captured in one hipGraph.
 * `*cur` is the level holding the current field (i2 in FDTDmain); it is
 * updated on return.  Per-step values at the source point (what
 * FDTDUtils::output prints) and at the receivers are recorded on the device;
 * read them with phc_fd_get_traces. */
phc_status phc_fd_run(phc_fd *h, int it0, int nsteps, int *cur);

/* Receivers: interior (i,j,k) triples; replaces the previous set. */
phc_status phc_fd_set_receivers(phc_fd *h, int nrec, const int *ijk);
/* Traces of the last phc_fd_run: src[nsteps], rec[nrec*nsteps] (receiver-major).
 * Either pointer may be NULL. */
phc_status phc_fd_get_traces(phc_fd *h, float *src, float *rec);

/* Field of a level as the reference holds it: interior+halo, IDX3_l order
 * ((nx+2lx)*(ny+2ly)*(nz+2lz) floats, FDTDmacros.hpp:14). */
phc_status phc_fd_get_field(phc_fd *h, int level, float *host);
phc_status phc_fd_set_field(phc_fd *h, int level, const float *host);
/* One value of a level at interior (i,j,k) (FDTDUtils::output's print). */
phc_status phc_fd_read_point(phc_fd *h, int level, int i, int j, int k, float *v);
/* FDTDUtils::write_io (FDTDutils.hpp:212-232): the x-z slice at j into host
 * [nz][nx] (k slowest, i fastest, as the reference writes it). */
phc_status phc_fd_get_slice_xz(phc_fd *h, int level, int j, float *host);
/* PHC_OK if every interior value of the level is finite, else PHC_ERR_NONFINITE. */
phc_status phc_fd_check_finite(phc_fd *h, int level);

/* Timing of the dominant kernel: mean duration in ms of the stencil launches
 * of the last phc_fd_run, measured with HIP events on the handle's stream. */
phc_status phc_fd_last_run_kernel_ms(phc_fd *h, double *ms_per_step);
/* Enable/disable the per-launch event timing used above (default off). */
phc_status phc_fd_set_timing(phc_fd *h, int enable);
/* Wait for all work queued on the handle's stream. */
phc_status phc_fd_synchronize(phc_fd *h);

/* Autotuner (no reference counterpart; the reference has fixed launch
 * shapes, FDTDmacros.hpp:57-61): times each block shape and chunk length on
 * scratch copies of the levels (the state is not advanced), keeps the fastest
 * and reports it.  cache_path (may be NULL) is a text cache keyed by grid,
 * radius and arithmetic; a cached entry is used without timing. */
phc_status phc_fd_autotune(phc_fd *h, int steps_per_trial, const char *cache_path, int *tile_j,
                           int *block_waves, int *chunk_i);

#ifdef __cplusplus
}
#endif
#endif
