// fd_tile.hpp -- the marching tile body of the FD step (inner3D + applySponge
// + fused source injection, src/fd/utilsFDTD/FDTDkernels.hpp:8-34, 108-128),
// shared by the single-grid step kernel (fd.hip) and the batched multi-shot
// kernel (batch.hip).
#pragma once

#include "fd_kernels.hpp"
#include "march.hpp"
#include "phc.h"

namespace phc {

// New value of one point, reference rounding: float lap accumulators, the
// update evaluated in double exactly as FDTDkernels.hpp:30-31 writes it, then
// the sponge product in float (FDTDkernels.hpp:114).
__device__ __forceinline__ float fd_update_ref(float pc, float pm, float v, double coef0,
                                               float lapx, float lapy, float lapz, float s)
{
#pragma clang fp contract(off)
    const double dpc = (double)pc;
    const double a = 2.0 * dpc - (double)pm;
    const double b = (((coef0 * dpc + (double)lapx) + (double)lapy) + (double)lapz);
    const float r = (float)(a + (double)v * b);
    return r * s;
}

// pml3D (FDTDkernels.hpp:36-73) at one point, every operation rounded where
// the reference's expression rounds.  e = eta at the point; the three
// first-difference terms use eta one step each way along their axis:
//   tg = (d eta/dx)(dp/dx) hd2_x + (d eta/dy)(dp/dy) hd2_y + (d eta/dz)(dp/dz) hd2_z
// ARITH_FP32 takes the eta differences as exp_ - exm etc. directly.
template <int ARITH>
__device__ __forceinline__ void fd_pml_terms(float exm, float ex0, float exp_, float eym, float ey0, float eyp,
                                             float ezm, float ez0, float ezp, float pxm, float pxp, float pym,
                                             float pyp, float pzm, float pzp, const float *hd2, float &e, float &tg)
{
#pragma clang fp contract(off)
    e = (ex0 + ey0) + ez0;
    if constexpr (ARITH == PHC_FD_ARITH_REF) {
        const float t1 = (((exp_ + ey0) + ez0) - ((exm + ey0) + ez0)) * (pxp - pxm) * hd2[0];
        const float t2 = (((ex0 + eyp) + ez0) - ((ex0 + eym) + ez0)) * (pyp - pym) * hd2[1];
        const float t3 = (((ex0 + ey0) + ezp) - ((ex0 + ey0) + ezm)) * (pzp - pzm) * hd2[2];
        tg = (t1 + t2) + t3;
    } else {
        tg = (exp_ - exm) * (pxp - pxm) * hd2[0] + (eyp - eym) * (pyp - pym) * hd2[1] +
             (ezp - ezm) * (pzp - pzm) * hd2[2];
    }
}

// New value and new phi of pml3D from the stencil's per-axis float sums.
// ARITH_REF: lap formed in double and stored as float, the update in double,
// phi in float with a double division (the reference's rounding).
// ARITH_FP32: lap32 is the whole Laplacian (the plain step's FMA chain); the
// two divisions as products with the hardware reciprocal (1 ulp), a tenth of
// the instructions of an IEEE division.
template <int ARITH>
__device__ __forceinline__ void fd_pml_update(float pc, float pm, float v, float ph, double coef0,
                                              float lapx, float lapy, float lapz, float lap32, float e, float tg,
                                              float &val, float &phin)
{
#pragma clang fp contract(off)
    if constexpr (ARITH == PHC_FD_ARITH_REF) {
        const double dpc = (double)pc;
        const float lap = (float)(((coef0 * dpc + (double)lapx) + (double)lapy) + (double)lapz);
        const double de = (double)e;
        const double num = (((2.0 - (double)(e * e)) + 2.0 * de) * dpc - (double)pm) + (double)(v * (ph + lap));
        val = (float)(num / (1.0 + 2.0 * de));
        phin = (float)((double)(ph - tg) / (1.0 + de));
    } else {
        val = ((2.f - e * e + 2.f * e) * pc - pm + v * (ph + lap32)) * __builtin_amdgcn_rcpf(1.f + 2.f * e);
        phin = (ph - tg) * __builtin_amdgcn_rcpf(1.f + e);
    }
}

// Face damping of one axis at index i of n (FDTDinit.hpp:343-395: the high
// face loop runs last, so it wins where the faces overlap).
__device__ __forceinline__ float fd_face_damping(const FdStepArgs &a, int i, int n)
{
    const int d = (i >= n - a.sw) ? a.sw - (n - i) : ((i < a.sw) ? a.sw - i : -1);
    return d >= 0 ? a.stab[d] : 1.f;
}

// One launch = inner3D + applySponge over the grid.
//
// A block owns a 64(k) x TJ(j) column and marches along i over chunk_i planes.
// Per thread and point: a ring of 2R+2 registers holds pn along i (planes
// i-R..i+R plus the one in flight); a ring of R+1 registers per halo element
// holds the j/k halo of the planes staged next; two register sets hold pnm1
// and vp^2dt^2 of planes i and i+1.  The march is unrolled by 2R+2, so every
// ring index is a compile-time constant and no value is ever moved between
// registers: a load has a full plane of work to complete before its register
// is read.  The halo of plane p is requested together with pn(p) of the
// register queue, i.e. when the neighbouring tile requests the same lines as
// its interior, so halo reads hit L2.  The current plane plus its halo is
// staged in LDS (double buffer, one LDS-only barrier per plane).
// IT is the offset type: 32-bit byte offsets when every array is < 4 GiB.
template <int R, int TY, int NJ>
struct FdTileShape {
    static constexpr int TJ = TY * NJ;
    static constexpr int LW = FD_TK + 2 * R;
    static constexpr int LH = TJ + 2 * R;
};

// March of one tile column (k-tile kt, j-tile jt) of one grid over planes [i0, i1).
// PML: points outside the PML's inner box take the pml3D update (the
// stencil values come from the same register ring and LDS tile), so one
// launch covers the slabs and the inner box.
template <int R, int TY, int NJ, int ARITH, typename IT, bool PML = false, class G>
__device__ __forceinline__ void fd_tile(const FdStepArgs &a, const G &g, int kt, int jt, int i0, int i1,
                                        float (*tile)[FdTileShape<R, TY, NJ>::LH][FdTileShape<R, TY, NJ>::LW])
{
    constexpr int TJ = TY * NJ;
    constexpr int LW = FD_TK + 2 * R;
    constexpr int NH = (2 * R + TY - 1) / TY;  // j-halo rows per thread
    constexpr int Q = 2 * R + 2;                     // queue ring = unroll factor
    constexpr int F = R + 1;                         // halo ring (Q == 2F)
    constexpr IT S = sizeof(IT) == 4 ? 4 : 1;        // offset units per element

    const FdLayout &L = g.L;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int k0 = kt * FD_TK, j0 = jt * TJ;
    const int k = k0 + tx;
    const IT pitch = (IT)L.pitch_j * S;

    // Every load is unconditional: lanes or rows outside the stored grid read
    // the interior origin instead (their offsets point there), and what they
    // stage in LDS is never read by a point that is computed.
    const bool kload = k < L.nz + R;
    const bool kcomp = k >= g.klo && k < g.khi;
    int jr[NJ];
    bool jcomp[NJ], inJ[NJ];
    float fyv[NJ];
    IT colq[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        jr[jj] = j0 + ty * NJ + jj;
        const bool jl = kload && (jr[jj] < L.ny + R);
        jcomp[jj] = kcomp && (jr[jj] >= g.jlo) && (jr[jj] < g.jhi);
        inJ[jj] = (jr[jj] >= a.sw) && (jr[jj] < L.ny - a.sw);
        fyv[jj] = (a.sw > 0 && jr[jj] < L.ny) ? g.fy[jr[jj]] : 1.f;
        colq[jj] = (IT)fd_off(L, 0, jl ? jr[jj] : 0, jl ? k : 0) * S;
    }
    const bool inK = (k >= a.sw) && (k < L.nz - a.sw);
    const float fzv = (a.sw > 0 && kcomp) ? g.fz[k] : 1.f;
    // PML profiles along k (per lane) and j (per row), fixed during the march
    float ezm = 0.f, ez0 = 0.f, ezp = 0.f;
    float eym[NJ], ey0[NJ], eyp[NJ];
    bool pslab_jk[NJ];
    if constexpr (PML) {
        if (kcomp) {
            ezm = a.pml_ez[k];
            ez0 = a.pml_ez[k + 1];
            ezp = a.pml_ez[k + 2];
        }
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            const bool jv = jcomp[jj];
            eym[jj] = jv ? a.pml_ey[jr[jj]] : 0.f;
            ey0[jj] = jv ? a.pml_ey[jr[jj] + 1] : 0.f;
            eyp[jj] = jv ? a.pml_ey[jr[jj] + 2] : 0.f;
            pslab_jk[jj] = jr[jj] < a.py3 || jr[jj] >= a.py4 || k < a.pz3 || k >= a.pz4;
        }
    }
    const bool khalo = tx < 2 * R;
    const int kk = (tx < R) ? (k0 - R + tx) : (k0 + FD_TK + tx - R);
    const int kcol = kk - k0 + R;
    IT colk[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const bool l = khalo && kk < L.nz + R && jr[jj] < L.ny + R;
        colk[jj] = (IT)fd_off(L, 0, l ? jr[jj] : 0, l ? kk : 0) * S;
    }
    int jhrow[NH];
    IT colh[NH];
#pragma unroll
    for (int m = 0; m < NH; ++m) {
        const int h = ty + TY * m;
        const int jh = (h < R) ? (j0 - R + h) : (j0 + TJ + h - R);
        jhrow[m] = jh - j0 + R;
        const bool l = (h < 2 * R) && (jh < L.ny + R) && kload;
        colh[m] = (IT)fd_off(L, 0, l ? jh : 0, l ? k : 0) * S;
    }

    float q[NJ][Q];                  // pn of plane p at slot (p - i0 + R) % Q
    float fk[NJ][F], fj[NH][F];      // halo of plane p at slot (p - i0) % F
    float pm[2][NJ], vv[2][NJ];      // pnm1 / vp of plane p in set (p - i0) % 2
    float ph[2][NJ];                 // PML: phi of plane p (slab points) in set (p - i0) % 2
    float ex[2][3];                  // PML: eta_x at p-1, p, p+1 (pml_ex[p..p+2]) in set (p - i0) % 2

    auto load_halo = [&](int slot, int p) {
        const IT po = (IT)p * pitch;
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) fk[jj][slot] = ldo(g.pn, colk[jj] + po);
#pragma unroll
        for (int m = 0; m < NH; ++m) fj[m][slot] = ldo(g.pn, colh[m] + po);
    };
    auto load_aux = [&](int set, int p) {
        const IT po = (IT)p * pitch;
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            pm[set][jj] = ldo_nt(g.pnm1, colq[jj] + po);
            vv[set][jj] = ldo_nt(g.vp, colq[jj] + po);
        }
        if constexpr (PML) {
            // phi only where plane p's point is in a slab; other points read
            // element 0, so the load needs no branch.  A cached load: as a
            // non-temporal one, every wavefront's read of that one line went
            // to L2 (2% of the PML step at 512^3)
            const bool sl = p < a.px3 || p >= a.px4;
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj)
                ph[set][jj] = ldo(a.pml_phi, (sl || pslab_jk[jj]) ? colq[jj] + po : (IT)0);
            ex[set][0] = a.pml_ex[p];
            ex[set][1] = a.pml_ex[p + 1];
            ex[set][2] = a.pml_ex[p + 2];
        }
    };
    auto stage = [&](int buf, int qslot, int hslot) {
        float(*tl)[LW] = tile[buf];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            tl[ty * NJ + jj + R][tx + R] = q[jj][qslot];
            if (khalo) tl[ty * NJ + jj + R][kcol] = fk[jj][hslot];
        }
#pragma unroll
        for (int m = 0; m < NH; ++m)
            if (ty + TY * m < 2 * R) tl[jhrow[m]][tx + R] = fj[m][hslot];
    };

    // prologue: queue planes i0-R .. i0+R (slots 0..2R), halo planes
    // i0 .. i0+R (slots 0..R), pnm1/vp of plane i0 (set 0); stage plane i0
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj)
#pragma unroll
        for (int l = 0; l <= 2 * R; ++l) q[jj][l] = ldo(g.pn, colq[jj] + (IT)(i0 - R + l) * pitch);
#pragma unroll
    for (int d = 0; d <= R; ++d) load_halo(d, i0 + d);
    load_aux(0, i0);
    stage(0, R, 0);
    lds_barrier();

    for (int ib = i0; ib < i1; ib += Q) {
#pragma unroll
        for (int u = 0; u < Q; ++u) {
            const int i = ib + u;
            if (i >= i1) break;
            // Prefetch for plane i+1, issued unconditionally (past the end of
            // the grid the plane index is clamped; those values are never
            // used): with a conditional prefetch the wait-count insertion has
            // to assume no loads were issued and drains every load in flight.
            {
                const int pq = min(i + 1 + R, L.nx + R - 1);
                const IT po = (IT)pq * pitch;
#pragma unroll
                for (int jj = 0; jj < NJ; ++jj) q[jj][(u + 1 + 2 * R) % Q] = ldo(g.pn, colq[jj] + po);
                load_halo((u + 1 + R) % F, pq);
                load_aux((u + 1) & 1, min(i + 1, L.nx - 1));
            }
            // compute plane i
            const float(*tl)[LW] = tile[u & 1];
            const bool inI = (i >= a.sw) && (i < L.nx - a.sw);
            const float fxv = (a.sw > 0) ? fd_face_damping(a, i, L.nx) : 1.f;
            const IT pi = (IT)i * pitch;
            float exm = 0.f, ex0 = 0.f, exp_ = 0.f;
            bool pslab_i = false;
            if constexpr (PML) {
                exm = ex[u & 1][0];
                ex0 = ex[u & 1][1];
                exp_ = ex[u & 1][2];
                pslab_i = i < a.px3 || i >= a.px4;
            }
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj) {
                if (!jcomp[jj]) continue;
                const int r = ty * NJ + jj + R;
                const int c = tx + R;
                const float pc = q[jj][(u + R) % Q];
                const float pmj = pm[u & 1][jj], vvj = vv[u & 1][jj];
                // stencil sums, once per point whichever update follows:
                // ARITH_REF per axis with the reference's rounding, ARITH_FP32
                // one FMA chain
                float lapx = 0.f, lapy = 0.f, lapz = 0.f, lap = 0.f;
                if constexpr (ARITH == PHC_FD_ARITH_REF) {
#pragma clang fp contract(off)
#pragma unroll
                    for (int l = 1; l <= R; ++l)
                        lapx += a.cx[l] * (q[jj][(u + R + l) % Q] + q[jj][(u + R - l) % Q]);
#pragma unroll
                    for (int l = 1; l <= R; ++l) lapy += a.cy[l] * (tl[r + l][c] + tl[r - l][c]);
#pragma unroll
                    for (int l = 1; l <= R; ++l) lapz += a.cz[l] * (tl[r][c + l] + tl[r][c - l]);
                } else {
                    lap = a.c0f * pc;
#pragma unroll
                    for (int l = 1; l <= R; ++l) {
                        lap = fmaf(a.cx[l], q[jj][(u + R + l) % Q] + q[jj][(u + R - l) % Q], lap);
                        lap = fmaf(a.cy[l], tl[r + l][c] + tl[r - l][c], lap);
                        lap = fmaf(a.cz[l], tl[r][c + l] + tl[r][c - l], lap);
                    }
                }
                if constexpr (PML) {
                    if (pslab_i || pslab_jk[jj]) {
                        const IT o = colq[jj] + pi;
                        float e, tg, val, phin;
                        fd_pml_terms<ARITH>(exm, ex0, exp_, eym[jj], ey0[jj], eyp[jj], ezm, ez0, ezp,
                                            q[jj][(u + R - 1) % Q], q[jj][(u + R + 1) % Q], tl[r - 1][c],
                                            tl[r + 1][c], tl[r][c - 1], tl[r][c + 1], a.pml_hd2, e, tg);
                        fd_pml_update<ARITH>(pc, pmj, vvj, ph[u & 1][jj], a.coef0, lapx, lapy, lapz, lap, e, tg,
                                             val, phin);
                        sto_nt(a.pml_phi, o, phin);
                        if ((int64_t)(o / S) == g.src_off) {
                            if (g.trace) *g.trace = val;
                            if (g.inj) val = add_nc(val, mul_nc(vvj, g.rhs_next));
                        }
                        sto_nt(g.pnm1, o, val);
                        continue;
                    }
                }
                float s = 1.f;
                if (a.sw > 0) {
                    if (inJ[jj] && inK && !inI) s = fxv;
                    if (inK && inI && !inJ[jj]) s = fyv[jj];
                    if (inJ[jj] && inI && !inK) s = fzv;
                }
                float val;
                if constexpr (ARITH == PHC_FD_ARITH_REF) {
                    // deferred damping of the previous level (x1 is exact)
                    const float pmv = mul_nc(pmj, a.damp_prev ? s : 1.f);
                    val = fd_update_ref(pc, pmv, vvj, a.coef0, lapx, lapy, lapz, s);
                } else {
                    const float pmv = pmj * (a.damp_prev ? s : 1.f);
                    val = s * fmaf(vvj, lap, 2.f * pc - pmv);
                }
                const IT o = colq[jj] + pi;
                if ((int64_t)(o / S) == g.src_off) {
                    if (g.trace) *g.trace = val;
                    if (g.inj) val = add_nc(val, mul_nc(vvj, g.rhs_next));
                }
                sto_nt(g.pnm1, o, val);
            }
            stage((u + 1) & 1, (u + 1 + R) % Q, (u + 1) % F);
            lds_barrier();
        }
    }
}

// The remainder slabs of one grid, one point per thread: thread t of the call
// updates point first + t of the k slab [g.krem, khi) x [jlo, jhi) x
// [ilo, ihi) (when g.krem > 0), then of the j slab [klo, krem or khi) x
// [g.jrem, jhi) x [ilo, ihi) (when g.jrem > 0), k fastest; the tiles cover
// the rest.  The same update as fd_tile, with the same operations in the
// same order, so a point gets the same bits from either path; the stencil
// values come straight from memory (a slab is at most 32 points wide along
// k or a quarter tile along j: these loads are L2 hits of the rows the
// neighbouring tiles stream).
template <int R, int ARITH, typename IT, bool PML = false, class G>
__device__ __forceinline__ void fd_rem_points(const FdStepArgs &a, const G &g, int64_t first, int nthreads)
{
    constexpr IT S = sizeof(IT) == 4 ? 4 : 1;
    const FdLayout &L = g.L;
    (void)nthreads;
    int64_t t = first + (int64_t)threadIdx.y * blockDim.x + threadIdx.x;
    const int NI = g.ihi - g.ilo;
    const int64_t nA = g.krem > 0 ? (int64_t)(g.khi - g.krem) * (g.jhi - g.jlo) * NI : 0;
    int i, j, k;
    if (t < nA) {
        const int W = g.khi - g.krem, NJ = g.jhi - g.jlo;
        k = g.krem + (int)(t % W);
        const int64_t rest = t / W;
        j = g.jlo + (int)(rest % NJ);
        i = g.ilo + (int)(rest / NJ);
    } else {
        t -= nA;
        if (g.jrem <= 0) return;
        const int kend = g.krem > 0 ? g.krem : g.khi;
        const int W = kend - g.klo, NJ = g.jhi - g.jrem;
        if (W <= 0 || NJ <= 0 || t >= (int64_t)W * NJ * NI) return;
        k = g.klo + (int)(t % W);
        const int64_t rest = t / W;
        j = g.jrem + (int)(rest % NJ);
        i = g.ilo + (int)(rest / NJ);
    }
    const IT si = (IT)L.pitch_j * S, sj = (IT)L.pitch_k * S;
    const IT o = (IT)fd_off(L, i, j, k) * S;
    const float pc = ldo(g.pn, o);
    const float pmj = ldo(g.pnm1, o), vvj = ldo(g.vp, o);
    float lapx = 0.f, lapy = 0.f, lapz = 0.f, lap = 0.f;
    if constexpr (ARITH == PHC_FD_ARITH_REF) {
#pragma clang fp contract(off)
#pragma unroll
        for (int l = 1; l <= R; ++l) lapx += a.cx[l] * (ldo(g.pn, o + l * si) + ldo(g.pn, o - l * si));
#pragma unroll
        for (int l = 1; l <= R; ++l) lapy += a.cy[l] * (ldo(g.pn, o + l * sj) + ldo(g.pn, o - l * sj));
#pragma unroll
        for (int l = 1; l <= R; ++l) lapz += a.cz[l] * (ldo(g.pn, o + l * S) + ldo(g.pn, o - l * S));
    } else {
        lap = a.c0f * pc;
#pragma unroll
        for (int l = 1; l <= R; ++l) {
            lap = fmaf(a.cx[l], ldo(g.pn, o + l * si) + ldo(g.pn, o - l * si), lap);
            lap = fmaf(a.cy[l], ldo(g.pn, o + l * sj) + ldo(g.pn, o - l * sj), lap);
            lap = fmaf(a.cz[l], ldo(g.pn, o + l * S) + ldo(g.pn, o - l * S), lap);
        }
    }
    float val;
    if constexpr (PML) {
        const bool slab = i < a.px3 || i >= a.px4 || j < a.py3 || j >= a.py4 || k < a.pz3 || k >= a.pz4;
        if (slab) {
            float e, tg, phin;
            fd_pml_terms<ARITH>(a.pml_ex[i], a.pml_ex[i + 1], a.pml_ex[i + 2], a.pml_ey[j], a.pml_ey[j + 1],
                                a.pml_ey[j + 2], a.pml_ez[k], a.pml_ez[k + 1], a.pml_ez[k + 2], ldo(g.pn, o - si),
                                ldo(g.pn, o + si), ldo(g.pn, o - sj), ldo(g.pn, o + sj), ldo(g.pn, o - S),
                                ldo(g.pn, o + S), a.pml_hd2, e, tg);
            fd_pml_update<ARITH>(pc, pmj, vvj, ldo(a.pml_phi, o), a.coef0, lapx, lapy, lapz, lap, e, tg, val, phin);
            sto_nt(a.pml_phi, o, phin);
            if ((int64_t)(o / S) == g.src_off) {
                if (g.trace) *g.trace = val;
                if (g.inj) val = add_nc(val, mul_nc(vvj, g.rhs_next));
            }
            sto_nt(g.pnm1, o, val);
            return;
        }
    }
    float s = 1.f;
    if (a.sw > 0) {
        const bool inI = (i >= a.sw) && (i < L.nx - a.sw), inJ = (j >= a.sw) && (j < L.ny - a.sw),
                   inK = (k >= a.sw) && (k < L.nz - a.sw);
        if (inJ && inK && !inI) s = fd_face_damping(a, i, L.nx);
        if (inK && inI && !inJ) s = g.fy[j];
        if (inJ && inI && !inK) s = g.fz[k];
    }
    if constexpr (ARITH == PHC_FD_ARITH_REF) {
        const float pmv = mul_nc(pmj, a.damp_prev ? s : 1.f);
        val = fd_update_ref(pc, pmv, vvj, a.coef0, lapx, lapy, lapz, s);
    } else {
        const float pmv = pmj * (a.damp_prev ? s : 1.f);
        val = s * fmaf(vvj, lap, 2.f * pc - pmv);
    }
    if ((int64_t)(o / S) == g.src_off) {
        if (g.trace) *g.trace = val;
        if (g.inj) val = add_nc(val, mul_nc(vvj, g.rhs_next));
    }
    sto_nt(g.pnm1, o, val);
}

} // namespace phc
