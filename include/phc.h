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
    int lx, ly, lz;   /* stencil half widths (halo widths), each 1..6 (FDTDinit.hpp:96-110) */
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
    int chunk_i;       /* planes per block along the marching axis i; -m
                        * (m = 1..4): per-XCD schedule, the axis cut into
                        * 8 >> (m-1) plane ranges, each XCD's tiles marching
                        * one range (the table phc_sem_schedule_table
                        * describes, with one-plane groups) */
    int index64;       /* 1: force 64-bit element offsets (default: 32-bit when
                        * every array is below 4 GiB) */
    int k_remainder;   /* a last, partial 64-point tile along k: 0 (default)
                        * updated point by point when it holds at most 32
                        * points, as a whole tile otherwise; likewise a last
                        * tile along j holding at most a quarter of tile_j
                        * rows; 1 always point by point; -1 always as a tile.
                        * Same values either way. */
} phc_fd_params;

phc_status phc_fd_create(const phc_fd_params *p, phc_fd **out);
phc_status phc_fd_destroy(phc_fd *h);

/* FDTDKernel::addRHS (FDTDkernels.hpp:121-128): level[cb][src] += vp[src]*rhs[it]. */
phc_status phc_fd_add_rhs(phc_fd *h, int it, int cb);
/* FDTDKernel::addRHS with the caller's sample (FDTDmain.cpp:59 passes
 * RHSTerm[itSample] through FDTDMODELS): level[cb](i,j,k) += vp(i,j,k) * rhs. */
phc_status phc_fd_add_rhs_value(phc_fd *h, int cb, int i, int j, int k, float rhs);
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

/* PML boundary (FDTDInit -usePML, FDTDinit.hpp:155-198, 263-275;
 * FDTDKernel::computeOneStepPML / pml3D, FDTDkernels.hpp:36-104): slabs of
 * ndamp[a] points at both ends of axis a run pml3D with
 * eta(i,j,k) = etax[i+1] + etay[j+1] + etaz[k+1] (profiles over -1..n, n+2
 * values, e.g. from phc_fd_pml_profile) and hd2[a] = 1/(4 d_a^2); the rest
 * runs inner3D; the auxiliary field phi starts at 0.  Needs sponge_width 0.
 * ndamp all 0 removes the PML. */
phc_status phc_fd_set_pml(phc_fd *h, const int *ndamp, const float *etax, const float *etay, const float *etaz,
                          const float *hd2);
/* FDTDKernel::computeOneStepPML (FDTDkernels.hpp:87-104): level ca <- new
 * level from cb (and ca), pml3D on the slabs, inner3D inside.  phc_fd_run
 * uses it when a PML is set. */
phc_status phc_fd_compute_one_step_pml(phc_fd *h, int ca, int cb);

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

/* ==================================================================== FD64
 * The FD time loop of FDTDmain.cpp:55-81 (addRHS + inner3D + applySponge,
 * FDTDkernels.hpp:8-34, 108-128) in FP64: fields, vp^2 dt^2, coefficients,
 * source samples and sponge factors (exp(alpha d^2) in double) all double.
 * The accuracy yardstick of the FP32 paths (half width 4; the arith and tiling
 * fields of phc_fd_params are ignored).  Synchronous: phc_fd64_run returns
 * when the steps are done.
 */
typedef struct phc_fd64 phc_fd64;
phc_status phc_fd64_create(const phc_fd_params *p, phc_fd64 **out);
phc_status phc_fd64_destroy(phc_fd64 *h);
phc_status phc_fd64_set_receivers(phc_fd64 *h, int nrec, const int *ijk);
/* Samples it0 .. it0+nsteps-1, continuing from the handle's current state. */
phc_status phc_fd64_run(phc_fd64 *h, int it0, int nsteps);
/* Last run: src[nsteps] at the source after each step, rec[nrec*nsteps]. */
phc_status phc_fd64_get_traces(phc_fd64 *h, double *src, double *rec);
/* Current (current != 0) or previous level, extended layout as phc_fd_get_field. */
phc_status phc_fd64_get_field(phc_fd64 *h, int current, double *host);

/* ===================================================================== SEM
 * Replaces SEMsolver (src/sem/solver/SEMsolver.hpp:18-91, SEMsolver.cpp:12-212)
 * on the structured hexahedral box mesh of SEMmesh (src/sem/mesh/SEMmesh.cpp).
 * Node numbering is the reference's: node = ix + iz*nx + iy*nx*nz
 * (SEMmesh.cpp:214-223); element e = ex_i + ez_i*ex + ey_i*ex*ez.
 */
typedef struct phc_sem phc_sem;

/* Step kernel (SEMsolver::computeOneStep, SEMsolver.cpp:21-125). */
#define PHC_SEM_SEPARABLE 0 /* assembled operator as a radius-p stencil (axis-aligned box) */
#define PHC_SEM_ELEMENT 1   /* element-by-element sum factorisation + atomic scatter */
/* element kernel with the scatter in 8 colour passes (elements of equal
 * (x, z, y) index parities share no node): plain adds in a fixed order, so
 * results repeat bit for bit.  The 3D form of SEM_MESHCOLOR
 * (SEMmesh.cpp:380-429, commonMacros.hpp:58-63), whose 4 colours cover
 * only the first element layer in 3D. */
#define PHC_SEM_ELEMENT_COLOR 2
/* element kernel that writes every node once: a block owns the nodes of a
 * tile of elements and marches the element layers; K_e u_e of the layer's
 * elements is formed in LDS (one lane per element node plane), every owned
 * node sums the contributions of the elements holding it in a fixed order,
 * and the leapfrog update follows (SEMsolver.cpp:46-89 in one pass, no y
 * array) */
#define PHC_SEM_ELEMENT_MARCH 3
/* the colour-scatter element kernel with its contractions (gradient and
 * divergence, SEMQkGLIntegrals.hpp:363-520) on the matrix cores: one element
 * per wavefront, each direction's contraction as FP32 16x16x4 MFMA products
 * (lines of the element x the padded 1D derivative matrix).  Orders 1..5. */
#define PHC_SEM_ELEMENT_MFMA 4

typedef struct {
    int ex, ey, ez;        /* elements per axis (SEMproxy.cpp:12-14) */
    float lx, ly, lz;      /* domain lengths (SEMproxy.cpp:16-18) */
    int order;             /* GLL order p, 1..5 (SEMdata.hpp:14) */
    const float *model;    /* wave speed per element (SEMmesh::getModel) */
    float dt;              /* time step (SEMdata.hpp:15) */
    const float *rhs;      /* source wavelet, nsamples values (SEMproxy.cpp:108-115) */
    int nsamples;
    int src_element;       /* element holding the source (SEMproxy.cpp:101-106) */
    int kernel;            /* PHC_SEM_* */
    int weights_ref;       /* 1: the reference's 8-decimal GLL weights; 0: exact */
    int tile_j, block_waves, chunk_i; /* separable kernel tiling (0: default); chunk_i = -m,
                                         m = 1..4: per-XCD schedule of 8 >> (m-1) plane
                                         ranges x 2^(m-1) column bands */
    int index64;           /* 1: 64-bit offsets in the separable and 2D kernels even
                              when every array is < 4 GiB (testing) */
    int k_remainder;       /* separable kernel: a last, partial 64-node tile along x
                              updated node by node when it holds at most 32 nodes (0,
                              default), always (1), or marched as a tile (-1) */
} phc_sem_params;

phc_status phc_sem_create(const phc_sem_params *p, phc_sem **out);
phc_status phc_sem_destroy(phc_sem *h);
/* Number of nodes (nx*ny*nz) and of nodes per axis. */
phc_status phc_sem_sizes(phc_sem *h, int *num_nodes, int *nx, int *ny, int *nz);
/* SEMsolver::computeOneStep (SEMsolver.cpp:21-125): inject the source sample
 * timeSample into level i2 at the first node of the source element, then
 * level i1 <- 2 level i2 - level i1 - dt^2 (K level i2)/M on interior nodes. */
phc_status phc_sem_compute_one_step(phc_sem *h, int time_sample, int i1, int i2);
/* SEMsolver::computeOneStep with the caller's sources (SEMproxy.cpp:48 passes
 * RHSTerm(:, timeSample) and rhsElement, SEMsolver.cpp:37-42): for r = 0..nrhs-1
 * level i2 at the first node of rhs_element[r] += dt^2 c^2 rhs_value[r], then
 * the step. */
phc_status phc_sem_compute_one_step_rhs(phc_sem *h, int i1, int i2, int nrhs, const int *rhs_element,
                                        const float *rhs_value);
/* Source of the fused loop (phc_sem_run): element and samples, replacing the
 * ones given at create (SEMproxy::init_source, SEMproxy.cpp:93-119). */
phc_status phc_sem_set_source(phc_sem *h, int src_element, int nsamples, const float *rhs);
/* SEMproxy::run (SEMproxy.cpp:41-64) for steps it0..it0+nsteps-1, injection
 * fused into the preceding step, captured in one hipGraph.  *i1 / *i2 are the
 * reference's level indices and are swapped as the reference swaps them. */
phc_status phc_sem_run(phc_sem *h, int it0, int nsteps, int *i1, int *i2);
/* Receivers at node ids; traces of the last phc_sem_run (receiver-major) and
 * the source-node trace (what SEMsolver::outputPnValues prints). */
phc_status phc_sem_set_receivers(phc_sem *h, int nrec, const int *nodes);
phc_status phc_sem_get_traces(phc_sem *h, float *src, float *rec);
/* Level as the reference holds it: pnGlobal(:, level) in node order. */
phc_status phc_sem_get_field(phc_sem *h, int level, float *host);
phc_status phc_sem_set_field(phc_sem *h, int level, const float *host);
/* Assembled lumped mass M (sum over elements of w w w |J| / c^2, FP64 then
 * rounded) in node order. */
phc_status phc_sem_get_mass(phc_sem *h, float *host);
/* y = K p for a host vector p (node order), through the handle's kernel. */
phc_status phc_sem_apply_stiffness(phc_sem *h, const float *p, float *y);
phc_status phc_sem_check_finite(phc_sem *h, int level);
phc_status phc_sem_set_timing(phc_sem *h, int enable);
phc_status phc_sem_last_run_kernel_ms(phc_sem *h, double *ms_per_step);
phc_status phc_sem_synchronize(phc_sem *h);
/* Per-XCD schedule of the separable kernel (no reference counterpart; host
 * only, for inspection and tests): for a march axis of march_planes planes,
 * tile_columns tile columns, cus CUs and split m (chunk_i = -m), the table
 * hardware block b reads: entries[3b..3b+2] = tile column, first plane, end
 * plane (end <= first: no work).  With entries NULL only *n_entries is set;
 * otherwise *n_entries is the room in entries (in triples) on entry. */
phc_status phc_sem_schedule_table(int march_planes, int tile_columns, int cus, int split, int *entries,
                                  int *n_entries);
/* Autotuner of the separable kernel (no reference counterpart): block shape
 * and chunk length (or -1..-4, the per-XCD schedules) timed on scratch copies
 * of the levels; optional text cache.
 * A no-op returning zeros for the element kernel. */
phc_status phc_sem_autotune(phc_sem *h, int steps_per_trial, const char *cache_path, int *tile_j,
                            int *block_waves, int *chunk_i);
/* ==================================================================== SEM64
 * SEMsolver::computeOneStep + the SEMproxy loop (SEMsolver.cpp:21-125,
 * SEMproxy.cpp:41-119) in FP64 on the structured box: fields, lumped mass,
 * dt^2 W/M and the assembled 1D operators all double, source samples double.
 * The accuracy yardstick of the FP32 SEM paths (the kernel and tiling fields
 * of phc_sem_params are ignored; the source must be on an interior node).
 * Synchronous.  Levels start as i1 = 0, i2 = 1 (SEMproxy.cpp:41-45).
 */
typedef struct phc_sem64 phc_sem64;
phc_status phc_sem64_create(const phc_sem_params *p, phc_sem64 **out);
phc_status phc_sem64_destroy(phc_sem64 *h);
phc_status phc_sem64_set_receivers(phc_sem64 *h, int nrec, const int *nodes);
phc_status phc_sem64_run(phc_sem64 *h, int it0, int nsteps);
/* Last run: src[nsteps] = source node after each step, rec[nrec*nsteps]. */
phc_status phc_sem64_get_traces(phc_sem64 *h, double *src, double *rec);
/* current != 0: the field after the last step (level i2), else level i1; node order. */
phc_status phc_sem64_get_field(phc_sem64 *h, int current, double *host);

/* ================================================================= SEM2D
 * SEMsolver::computeOneStep of the reference built with SEM2D
 * (src/sem/solver/SEMsolver.cpp:21-125 including the SEM2D block 91-123):
 * ex x ez quadrilaterals of order p on the x-z plane (SEMmesh with ey = 0,
 * SEMmesh.cpp:3-26), node = ix + iz*nx, element e = i + k*ex; leapfrog on the
 * interior nodes and the first-order paraxial absorbing condition on the four
 * edges.  Parameters are phc_sem_params with ey = 0 (ly ignored, kernel
 * PHC_SEM_SEPARABLE, chunk_i = rows per wavefront, 0 = default).
 */
typedef struct phc_sem2d phc_sem2d;
phc_status phc_sem2d_create(const phc_sem_params *p, phc_sem2d **out);
phc_status phc_sem2d_destroy(phc_sem2d *h);
phc_status phc_sem2d_sizes(phc_sem2d *h, int *num_nodes, int *nx, int *nz);
/* inject sample time_sample into level i2 at the first node of the source
 * element (SEMsolver.cpp:39-43), then level i1 <- the new level */
phc_status phc_sem2d_compute_one_step(phc_sem2d *h, int time_sample, int i1, int i2);
/* the same with the caller's sources (SEMsolver.cpp:37-42), as phc_sem_compute_one_step_rhs */
phc_status phc_sem2d_compute_one_step_rhs(phc_sem2d *h, int i1, int i2, int nrhs, const int *rhs_element,
                                          const float *rhs_value);
/* source of the fused loop, as phc_sem_set_source */
phc_status phc_sem2d_set_source(phc_sem2d *h, int src_element, int nsamples, const float *rhs);
/* SEMproxy::run (SEMproxy.cpp:41-64) for steps it0..it0+nsteps-1, captured
 * in one hipGraph; *i1 / *i2 swapped as the reference swaps them */
phc_status phc_sem2d_run(phc_sem2d *h, int it0, int nsteps, int *i1, int *i2);
phc_status phc_sem2d_set_receivers(phc_sem2d *h, int nrec, const int *nodes);
/* last run: src[nsteps] = source node after each step (what
 * SEMsolver::outputPnValues prints), rec[nrec*nsteps] */
phc_status phc_sem2d_get_traces(phc_sem2d *h, float *src, float *rec);
phc_status phc_sem2d_get_field(phc_sem2d *h, int level, float *host);
phc_status phc_sem2d_set_field(phc_sem2d *h, int level, const float *host);
/* lumped mass M (with 1/c^2) and boundary term S (sum over boundary faces of
 * w (h/2) / c, the ShGlobal of SEMsolver.cpp:95-110), per node in node order
 * (S is 0 inside); FP64 sums rounded once.  Either pointer may be NULL. */
phc_status phc_sem2d_get_mass(phc_sem2d *h, float *mass, float *boundary);
phc_status phc_sem2d_check_finite(phc_sem2d *h, int level);
phc_status phc_sem2d_set_timing(phc_sem2d *h, int enable);
phc_status phc_sem2d_last_run_kernel_ms(phc_sem2d *h, double *ms_per_step);
phc_status phc_sem2d_synchronize(phc_sem2d *h);
/* Autotuner (no reference counterpart): rows per wavefront, timed on scratch
 * copies of the levels. */
phc_status phc_sem2d_autotune(phc_sem2d *h, int steps_per_trial, int *chunk);

/* SEMmesh::getElementNumberFromPoints (SEMmesh.cpp:266-285). */
int phc_sem_element_from_point(int ex, int ey, int ez, float lx, float ly, float lz, float x,
                               float y, float z);

/* ================================================================= BATCH
 * Many FD shots advanced together: one persistent launch per time step does
 * the FD step of every shot.  Each shot is the FD problem of the reference's
 * driver (src/fd/FDTDmain.cpp:9-108: FDTDKernel::addRHS + computeOneStepSB per
 * sample, FDTDkernels.hpp:75-128) on its own grid size, source position,
 * vp^2 dt^2 model and source samples; the stencil coefficients, sponge and
 * number of samples are shared.  The reference runs one shot per process;
 * these calls replace a loop of such processes.  Half width 4 only.
 */
typedef struct phc_batch phc_batch;

typedef struct {
    int nshots;
    const int *dims;          /* nshots x (nx, ny, nz) */
    const int *src;           /* nshots x (xs, ys, zs) */
    const float *const *vp;   /* per shot: nx*ny*nz values of vp^2 dt^2, IDX3 order */
    const float *const *rhs;  /* per shot: nsamples source samples */
    int nsamples;
    int half_width;           /* must be 4 */
    const float *coef;        /* half_width+1 stencil coefficients, all three axes */
    double coef0;             /* centre coefficient (FDTDinit.hpp:204-208) */
    int sponge_width;         /* 0..64 */
    float sponge_alpha;
    float init_value;         /* initial value of both levels */
    int arith;                /* PHC_FD_ARITH_* */
    int tile_j, block_waves;  /* 0 = default (16 x 16) */
    int chunk_i;              /* planes per work item, 0 = default (64); -m
                               * (m = 1..4): each shot's per-XCD schedule
                               * (as phc_fd_params.chunk_i), XCD x's entries
                               * in queue x, shot after shot */
    int blocks_per_cu;        /* 0 = default (1) */
    int k_remainder;          /* a shot's last, partial k tile: as in phc_fd_params */
} phc_batch_params;

phc_status phc_batch_create(const phc_batch_params *p, phc_batch **out);
phc_status phc_batch_destroy(phc_batch *h);
/* nrec receivers per shot at grid points ijk[(shot*nrec + r)*3 + 0..2]. */
phc_status phc_batch_set_receivers(phc_batch *h, int nrec, const int *ijk);
/* Samples it0 .. it0+nsteps-1 of every shot (FDTDmain.cpp:55-81 per shot),
 * captured in one hipGraph.  Shots continue from where the last run ended. */
phc_status phc_batch_run(phc_batch *h, int it0, int nsteps);
/* Traces of the last run: src[shot*nsteps + s] = new value at the shot's
 * source after step s (before the next sample is injected, FDTDutils.hpp:
 * 193-203); rec[(shot*nrec + r)*nsteps + s] likewise at each receiver.
 * Either pointer may be NULL. */
phc_status phc_batch_get_traces(phc_batch *h, float *src, float *rec);
/* One shot's current (current != 0) or previous level, halo included, in the
 * host layout of phc_fd_get_field. */
phc_status phc_batch_get_field(phc_batch *h, int shot, int current, float *host);
phc_status phc_batch_set_timing(phc_batch *h, int enable);
/* Mean duration in ms of one batched step launch of the last timed run. */
phc_status phc_batch_last_run_kernel_ms(phc_batch *h, double *ms_per_step);
phc_status phc_batch_synchronize(phc_batch *h);
/* Autotuner (no reference counterpart): block shape, chunk length and
 * resident blocks per CU timed on a copy of the shots (state not advanced). */
phc_status phc_batch_autotune(phc_batch *h, int steps_per_trial, int *tile_j, int *block_waves,
                              int *chunk_i, int *blocks_per_cu);

/* ======================================================= LEARNED BOUNDARY
 * A thin learned absorbing layer for the FD step, in place of the sponge
 * (FDTDinit.hpp:319-396, FDTDkernels.hpp:108-118) or the PML (FDTDkernels.hpp:
 * 36-104).  No reference counterpart.  The grid ends PHC_LB_WIDTH points
 * outside the region of interest, with no damping; at each of those points,
 * face f's network predicts the term of the radius-4 update that would read
 * beyond the grid edge:
 *   x = ((pn - h)/s at offsets -4..4 along the inward normal,
 *        (pnm1 - h)/s at the same offsets, sqrt(vp^2dt^2 c_1)),
 *   s = max |x_field - h| (h: the value the halo holds),
 *   new value += vp^2dt^2 c_1 s y,
 *   y = w3 . relu(W2 relu(W1 x + b1[depth]) + b2) + b3
 * with 19 inputs, two hidden layers of 32, bf16 weights and FP32 accumulation
 * (MFMA).  Parameters per face, FP32, in this order: W1[32][19], b1[4][32]
 * (one per layer depth 0..3 from the edge), W2[32][32], b2[32], w3[32], b3;
 * faces 2*axis+side (axis 0 = x/i, 1 = y/j, 2 = z/k; side 0 = low).
 * Training samples are PHC_LB_SAMPLE_FLOATS floats: x[19], target, depth,
 * index of the point in its face's layer (compact order below, as a float);
 * the target is the missing term over c_1 s, taken from a wide-sponge run.
 */
typedef struct phc_lb phc_lb;
#define PHC_LB_WIDTH 4
#define PHC_LB_FACES 6
#define PHC_LB_PARAMS_PER_FACE 1825
#define PHC_LB_SAMPLE_FLOATS 22

/* A model with He-uniform random weights (seeded). */
phc_status phc_lb_create(uint64_t seed, phc_lb **out);
phc_status phc_lb_destroy(phc_lb *h);
/* FP32 parameters, PHC_LB_FACES x PHC_LB_PARAMS_PER_FACE.  set resets Adam. */
phc_status phc_lb_get_params(phc_lb *h, float *host);
phc_status phc_lb_set_params(phc_lb *h, const float *host);
/* Training-sample store: capacity per face (clears it), clear, counts[6]. */
phc_status phc_lb_set_capacity(phc_lb *h, int64_t per_face);
phc_status phc_lb_clear_samples(phc_lb *h);
phc_status phc_lb_sample_counts(phc_lb *h, int64_t *counts);
phc_status phc_lb_get_samples(phc_lb *h, int face, int64_t first, int64_t n, float *host);
phc_status phc_lb_add_samples(phc_lb *h, int face, const float *host, int64_t n);
/* Append samples from an FD run with a wide sponge: the layer points of the
 * region origin[3] .. origin+dims of its grid, inputs from levels cur
 * (current) and 1-cur (previous), the targets from the grid beyond the
 * region.  halo: the value the thin grid's halo will hold.  Each point is kept
 * with probability keep, and only if its s >= s_min.  The region plus 4
 * points must lie outside the wide run's sponge. */
phc_status phc_lb_collect(phc_lb *h, phc_fd *wide, int cur, const int *origin, const int *dims, float halo,
                          float keep, float s_min, uint64_t seed);
/* Append on-policy samples: `thin` is a run on the region itself (grid =
 * region, typically with this model attached) and `wide` the same shot on the
 * region plus margin and sponge (the region at origin of its grid), both at
 * step n (current levels thin_cur, wide_cur).  Inputs come from the thin
 * levels (the states the network meets), the target from the wide field
 * beyond the region, over the thin s: sum_{l>d} (c_l/c_1)(W(P + l n_out) - h)
 * / s_thin; points whose |target| exceeds 16 are skipped. */
phc_status phc_lb_collect_onpolicy(phc_lb *h, phc_fd *thin, int thin_cur, phc_fd *wide, int wide_cur,
                                   const int *origin, float halo, float keep, float s_min, uint64_t seed);
/* Mean squared error over every stored sample and, if grad != NULL, its
 * gradient with respect to the FP32 parameters (at their bf16 values). */
phc_status phc_lb_eval(phc_lb *h, float *loss, float *grad);
/* Adam (lr, 0.9, 0.999, 1e-8) on minibatches of batch_per_face random samples
 * of every face; loss_history[s] (may be NULL) = minibatch loss at step s. */
phc_status phc_lb_train(phc_lb *h, int steps, int batch_per_face, float lr, uint64_t seed, float *loss_history);
/* No counterpart in the reference.  phc_lb_train for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same steps and the same values, with
 * the per-call buffers (minibatch indices, per-block partial gradients and
 * losses, loss history) in one static device array per GPU instead of
 * buffers allocated per call.  Limits: batch_per_face <= 8192 and
 * steps <= 16384 per call (PHC_ERR_INVALID beyond).  Calls that overlap on
 * one GPU share that array and race. */
phc_status phc_lb_train_noalloc(phc_lb *h, int steps, int batch_per_face, float lr, uint64_t seed,
                                float *loss_history);
/* Gaussian noise of standard deviation std added to the field inputs of every
 * training sample (the clean target kept), from the next phc_lb_train on:
 * the network then learns not to amplify errors in its own inputs, which a
 * rollout feeds back.  0 (the default) trains on the samples as stored. */
phc_status phc_lb_set_train_noise(phc_lb *h, float std);
/* Network output y of face `face` for n given samples (MFMA inference path). */
phc_status phc_lb_infer(phc_lb *h, int face, const float *samples, int64_t n, float *y);
/* Attach a copy of the model's bf16 weights to an FD handle (lb NULL detaches).
 * Needs half width 4, equal coefficients on the three axes, sponge_width 0,
 * sides 16..4000 and the source outside the layer; halo is the value the halo
 * holds (the handle's init_value).  Every step then runs the correction pass,
 * the stencil and the apply pass.  layer_damping (alpha in [-1, 0], 0 = none)
 * multiplies the new value at depth d of a face's layer by
 * exp(alpha (4-d)^2), once per face whose layer holds the point. */
phc_status phc_fd_set_learned_boundary(phc_fd *h, const phc_lb *lb, float halo, float layer_damping);
/* The corrections the next step would add (vp^2dt^2 c_1 s y per layer point)
 * from levels cur / 1-cur, compact per face: face f, depth d (from its edge),
 * u, v = the other two axes in increasing order, index ((d*n_u)+u)*n_v+v. */
phc_status phc_fd_lb_corrections(phc_fd *h, int cur, float *host);

#ifdef __cplusplus
}
#endif
#endif
