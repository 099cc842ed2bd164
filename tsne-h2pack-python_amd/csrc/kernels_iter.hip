// kernels_iter.hip -- per-iteration kernels of the t-SNE gradient path on gfx950:
//   * attractive CSR force with the gains/momentum update fused into its epilogue
//     (one wavefront per row),
//   * deterministic mean reduction for the re-centring,
//   * direct O(N^2) repulsive sums (FP32 product path, FP64 check path),
//   * Z from the repulsive sums, KL divergence,
//   * perplexity bisection (one wavefront per row, FP64).
//
// State layout in HBM (per workspace):
//   Y64, uY, gains   double [n][D]   interleaved, master copy of the embedding
//   Yg               Vec<D>  [n]     float32 gather copy (float2 in 2D, float4 in 3D)
//   F                float   [n][D]  unscaled repulsive sum  sum_j K_ij (y_i - y_j)
//   s                float   [n]     sum_{j != i} K_ij
//   scal             double  [..]    device scalars: Z, the pending centring mean, ...
// The row order of all of these is the caller's point order.
#include <algorithm>
#include <cstdlib>

#include "common.h"
#include "kernels.h"
#include "tiles.h"

namespace tsne {

// ---------------------------------------------------------------------------------
// Attractive force + fused update.
//
// Reference: attractive force src/tsne_h2pack.c:199-356 (2D float32 gather,
// src/tsne_attr_2d.h:22-67); update src/tsne_h2pack.c:1944-1968; centring :1978-1984.
//
// A group of 16 lanes per CSR row (4 rows per wavefront): lanes stride over the row's
// nonzeros, gather the neighbour from the float32 copy Yg and accumulate
// p_ij w_ij (y_i - y_j) in float32; a 16-lane reduction gives the row's force.  Lanes
// 0..D-1 of the group then apply the update to coordinate d of the row in FP64:
//     g      = 4*exag*attr_d - (4/Z) F_d            (F: reused or fresh repulsive sum)
//     gains  = (uY*g < 0) ? gains + 0.2 : gains * 0.8, floored at 0.01
//     uY     = mom*uY - lr*gains*g
//     y      = (Y64 - mean_prev) + uY
// and write Y64, uY, gains and the next gather copy Yg_next.  mean_prev is the mean
// removed lazily: the forces are translation invariant, so centring is applied when Y64
// is next read instead of in an extra pass.  Each block covers a fixed contiguous row
// range and writes the sum of its new coordinates in a fixed order, for the next mean.
// ---------------------------------------------------------------------------------

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// End of one row, shared by the attractive kernels: the group's LANES partial forces are
// reduced, lanes 0..D-1 apply the update to coordinate `sub` in FP64 and store Y64, uY,
// gains; lane 0 writes the row's next gather record.  Returns the new coordinate (lanes
// 0..D-1) for the centring sums, 0 elsewhere.
template <int D, int LANES>
__device__ __forceinline__ double row_epilogue(float ax, float ay, float az, int sub, size_t kd,
                                               double f_d, double u_d, double g_d, double y_d,
                                               double m_d, double exag4, double rep_scale,
                                               double mom, double lr, double *__restrict__ Y64,
                                               double *__restrict__ uY,
                                               double *__restrict__ gains,
                                               typename Vec<D>::T *__restrict__ rec_out)
{
    for (int off = LANES / 2; off > 0; off >>= 1) {
        ax += __shfl_xor(ax, off, 64);
        ay += __shfl_xor(ay, off, 64);
        if constexpr (D == 3) az += __shfl_xor(az, off, 64);
    }
    double ynew = 0.0;
    if (sub < D) {
        const float a = (sub == 0) ? ax : (sub == 1) ? ay : az;
        const double g = exag4 * (double)a - rep_scale * f_d;
        double u = u_d, gn = g_d;
        gn = (u * g < 0.0) ? gn + 0.2 : gn * 0.8;
        if (gn < 0.01) gn = 0.01;
        u = mom * u - lr * (gn * g);
        ynew = (y_d - m_d) + u;
        Y64[kd] = ynew;
        uY[kd] = u;
        gains[kd] = gn;
    }
    // the new coordinates sit in lanes 0..D-1 of the group: assemble the record
    const int base = threadIdx.x & ~(LANES - 1);
    const double n0 = __shfl(ynew, base, 64);
    const double n1 = __shfl(ynew, base + 1, 64);
    const double n2 = (D == 3) ? __shfl(ynew, base + 2, 64) : 0.0;
    if (sub == 0) {
        typename Vec<D>::T rec;
        rec.x = (float)n0;
        rec.y = (float)n1;
        if constexpr (D == 3) { rec.z = (float)n2; rec.w = 0.f; }
        *rec_out = rec;
    }
    return ynew;
}

// LANES lanes per row (256 / LANES rows in flight per block), BATCH nonzeros per lane
// per batch.  VEC: P's columns and values are read as aligned 4-element chunks (BATCH
// chunks per lane per batch) with non-temporal loads, so the 8 bytes per nonzero of the
// stream take a quarter of the load instructions and do not displace the gather records
// from the caches; elements of a chunk outside the row are masked.
// ATTR_ONLY: the attractive sums only, written to A (float [n][D], unscaled) for
// update_kernel; Y64, uY, gains, Yg_next and block_sums are left alone.  This form runs
// on a second stream next to an H2 rebuild (the rebuild and the attractive force of an
// iteration both read only the iteration's embedding).  Its launch may be smaller than
// the nvblk logical blocks (a multiple of 8 workgroups): each workgroup then walks the
// logical blocks b, b + G, ... (same XCD), leaving CUs to the rebuild's narrow launches;
// every row's sums are the same whichever workgroup computes them.
template <int D, int LANES, int BATCH, bool VEC, bool ATTR_ONLY = false>
__global__ void __launch_bounds__(256)
attr_update_kernel(int n, const int *__restrict__ row_ptr, const int *__restrict__ col_idx,
                   const float *__restrict__ pval, const typename Vec<D>::T *__restrict__ Yg,
                   const float *__restrict__ F, double *__restrict__ Y64,
                   double *__restrict__ uY, double *__restrict__ gains,
                   typename Vec<D>::T *__restrict__ Yg_next, const double *__restrict__ scal,
                   double exag4, double mom, double lr, int rows_per_block,
                   double *__restrict__ block_sums, float *__restrict__ A = nullptr,
                   int nvblk = 0)
{
    using V = typename Vec<D>::T;
    __shared__ double gsum[(256 / LANES)][3];
    const int sub = threadIdx.x & (LANES - 1);
    const int grp = threadIdx.x / LANES;
    const double rep_scale = ATTR_ONLY ? 0.0 : rep_scale_of(scal);
    const double m_d = (!ATTR_ONLY && sub < D) ? scal[SLOT_MEAN + sub] : 0.0;
    // the fused form runs exactly one logical block per workgroup (nv = the grid): its
    // gsum reduction at the end of a block is not fenced against a next block's writes
    const int nv = ATTR_ONLY ? nvblk : (int)gridDim.x;
    for (int vb = blockIdx.x; vb < nv; vb += gridDim.x) {
        // XCD-aware placement: workgroups are dealt round-robin to the 8 XCDs, so logical
        // block L = (b % 8) * (G / 8) + b / 8 gives each XCD one contiguous eighth of the
        // rows; with the points renumbered along the tree, that XCD's gathers stay in its
        // own L2.  (G, the number of logical blocks, is a multiple of 8.)
        const int lblk = (vb & 7) * (nv >> 3) + (vb >> 3);
        const int row0 = lblk * rows_per_block;
        const int row1 = min(n, row0 + rows_per_block);
        double my_sum = 0.0;  // lane d of the group: sum of new coordinate d over its rows

        for (int i = row0 + grp; i < row1; i += (256 / LANES)) {
            const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
            // update operands of this row, fetched now and used after the reduction
            const size_t kd = (size_t)i * D + (sub < D ? sub : 0);
            double f_d = 0.0, u_d = 0.0, g_d = 0.0, y_d = 0.0;
            if (!ATTR_ONLY && sub < D) {
                f_d = (double)F[kd];
                u_d = uY[kd];
                g_d = gains[kd];
                y_d = Y64[kd];
            }
            const V yi = Yg[i];
            float ax = 0.f, ay = 0.f, az = 0.f;
            if constexpr (VEC) {
                const int c1 = (e1 + 3) >> 2;  // chunks [e0/4, c1) cover the row
                for (int cb = (e0 >> 2) + sub; cb < c1; cb += BATCH * LANES) {
                    int jj[BATCH][4];
                    float pp[BATCH][4];
    #pragma unroll
                    for (int t = 0; t < BATCH; t++) {
                        const int c = cb + t * LANES;
                        v4i cc = {i, i, i, i};
                        v4f pv = {0.f, 0.f, 0.f, 0.f};
                        if (c < c1) {
                            cc = __builtin_nontemporal_load((const v4i *)col_idx + c);
                            pv = __builtin_nontemporal_load((const v4f *)pval + c);
                        }
                        const int e = 4 * c;
    #pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const bool ok = e + u >= e0 && e + u < e1;
                            jj[t][u] = ok ? cc[u] : i;
                            pp[t][u] = ok ? pv[u] : 0.f;
                        }
                    }
                    V yy[BATCH][4];
    #pragma unroll
                    for (int t = 0; t < BATCH; t++)
    #pragma unroll
                        for (int u = 0; u < 4; u++) yy[t][u] = Yg[jj[t][u]];
    #pragma unroll
                    for (int t = 0; t < BATCH; t++)
    #pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float dx = yi.x - yy[t][u].x, dy = yi.y - yy[t][u].y;
                            float r2 = fmaf(dx, dx, fmaf(dy, dy, 1.0f));
                            float dz = 0.f;
                            if constexpr (D == 3) { dz = yi.z - yy[t][u].z; r2 = fmaf(dz, dz, r2); }
                            const float q = pp[t][u] * fast_rcp(r2);
                            ax = fmaf(q, dx, ax);
                            ay = fmaf(q, dy, ay);
                            if constexpr (D == 3) az = fmaf(q, dz, az);
                        }
                }
            }
            // a batch is BATCH nonzeros per lane: all its column loads are issued, then
            // all its gathers, so a row costs two memory round trips instead of one per step
            for (int base = e0 + sub; !VEC && base < e1; base += BATCH * LANES) {
                int jj[BATCH];
                float pp[BATCH];
    #pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const int e = base + t * LANES;
                    const bool ok = e < e1;
                    jj[t] = ok ? col_idx[e] : i;
                    pp[t] = ok ? pval[e] : 0.f;
                }
                V yy[BATCH];
    #pragma unroll
                for (int t = 0; t < BATCH; t++) yy[t] = Yg[jj[t]];
    #pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const float dx = yi.x - yy[t].x, dy = yi.y - yy[t].y;
                    float r2 = fmaf(dx, dx, fmaf(dy, dy, 1.0f));
                    float dz = 0.f;
                    if constexpr (D == 3) { dz = yi.z - yy[t].z; r2 = fmaf(dz, dz, r2); }
                    const float q = pp[t] * fast_rcp(r2);
                    ax = fmaf(q, dx, ax);
                    ay = fmaf(q, dy, ay);
                    if constexpr (D == 3) az = fmaf(q, dz, az);
                }
            }
            if constexpr (ATTR_ONLY) {
                for (int off = LANES / 2; off > 0; off >>= 1) {
                    ax += __shfl_xor(ax, off, 64);
                    ay += __shfl_xor(ay, off, 64);
                    if constexpr (D == 3) az += __shfl_xor(az, off, 64);
                }
                if (sub < D) A[kd] = (sub == 0) ? ax : (sub == 1) ? ay : az;
            } else {
                my_sum += row_epilogue<D, LANES>(ax, ay, az, sub, kd, f_d, u_d, g_d, y_d, m_d, exag4,
                                                 rep_scale, mom, lr, Y64, uY, gains, Yg_next + i);
            }
        }
        if constexpr (!ATTR_ONLY) {
            // per-group sums (lane d holds coordinate d), then groups in order
            if (sub < D) gsum[grp][sub] = my_sum;
            __syncthreads();
            if (threadIdx.x < D) {
                double t = 0.0;
                for (int w = 0; w < (256 / LANES); w++) t += gsum[w][threadIdx.x];
                block_sums[(size_t)lblk * 3 + threadIdx.x] = t;
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Probe of the attractive pass's memory pattern (diagnostics, tsne_attr_probe): the same
// row mapping and chunked non-temporal P reads as attr_update_kernel, with the per-nonzero
// arithmetic reduced to a sum.  GATHER = false: only P streams (sum of p); GATHER = true:
// P plus the neighbour gathers (sum of p y_j).  One float per row and dimension is written
// to A.  The fused kernel's time against these two measures how close it is to the
// ceiling of its own access pattern, apart from the HBM peak.
// ---------------------------------------------------------------------------------
template <int D, bool GATHER>
__global__ void __launch_bounds__(256)
attr_probe_kernel(int n, const int *__restrict__ row_ptr, const int *__restrict__ col_idx,
                  const float *__restrict__ pval, const typename Vec<D>::T *__restrict__ Yg,
                  int rows_per_block, float *__restrict__ A)
{
    using V = typename Vec<D>::T;
    constexpr int LANES = 16, BATCH = 3;
    const int sub = threadIdx.x & (LANES - 1);
    const int grp = threadIdx.x / LANES;
    const int lblk = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int row0 = lblk * rows_per_block;
    const int row1 = min(n, row0 + rows_per_block);
    for (int i = row0 + grp; i < row1; i += (256 / LANES)) {
        const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
        float ax = 0.f, ay = 0.f, az = 0.f;
        const int c1 = (e1 + 3) >> 2;
        for (int cb = (e0 >> 2) + sub; cb < c1; cb += BATCH * LANES) {
            int jj[BATCH][4];
            float pp[BATCH][4];
#pragma unroll
            for (int t = 0; t < BATCH; t++) {
                const int c = cb + t * LANES;
                v4i cc = {i, i, i, i};
                v4f pv = {0.f, 0.f, 0.f, 0.f};
                if (c < c1) {
                    cc = __builtin_nontemporal_load((const v4i *)col_idx + c);
                    pv = __builtin_nontemporal_load((const v4f *)pval + c);
                }
                const int e = 4 * c;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const bool ok = e + u >= e0 && e + u < e1;
                    jj[t][u] = ok ? cc[u] : i;
                    pp[t][u] = ok ? pv[u] : 0.f;
                }
            }
#pragma unroll
            for (int t = 0; t < BATCH; t++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if constexpr (GATHER) {
                        const V y = Yg[jj[t][u]];
                        ax = fmaf(pp[t][u], y.x, ax);
                        ay = fmaf(pp[t][u], y.y, ay);
                        if constexpr (D == 3) az = fmaf(pp[t][u], y.z, az);
                    } else {
                        ax += pp[t][u];
                        ay += (float)jj[t][u];
                    }
                }
        }
        for (int off = LANES / 2; off > 0; off >>= 1) {
            ax += __shfl_xor(ax, off, 64);
            ay += __shfl_xor(ay, off, 64);
            if constexpr (D == 3) az += __shfl_xor(az, off, 64);
        }
        if (sub < D) A[(size_t)i * D + sub] = (sub == 0) ? ax : (sub == 1) ? ay : az;
    }
}

int launch_attr_probe(int D, int n, int mode, const int *row_ptr, const int *col_idx,
                      const float *pval, const void *Yg, float *A, int nblocks, hipStream_t st)
{
    const int rpb = (n + nblocks - 1) / nblocks;
    if (mode == 2) return launch_attr_only(D, n, row_ptr, col_idx, pval, Yg, A, nblocks, 0, st);
    if (D == 2) {
        if (mode == 1)
            hipLaunchKernelGGL((attr_probe_kernel<2, true>), dim3(nblocks), dim3(256), 0, st, n,
                               row_ptr, col_idx, pval, (const float2 *)Yg, rpb, A);
        else
            hipLaunchKernelGGL((attr_probe_kernel<2, false>), dim3(nblocks), dim3(256), 0, st, n,
                               row_ptr, col_idx, pval, (const float2 *)Yg, rpb, A);
    } else {
        if (mode == 1)
            hipLaunchKernelGGL((attr_probe_kernel<3, true>), dim3(nblocks), dim3(256), 0, st, n,
                               row_ptr, col_idx, pval, (const float4 *)Yg, rpb, A);
        else
            hipLaunchKernelGGL((attr_probe_kernel<3, false>), dim3(nblocks), dim3(256), 0, st, n,
                               row_ptr, col_idx, pval, (const float4 *)Yg, rpb, A);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Attractive force + fused update over row tiles (points numbered along the tree).
//
// One workgroup per tile of `tile_rows` consecutive rows.  The gather records of the
// tile's distinct neighbours (tcols[toff[b] .. toff[b+1]), sorted, so the staging reads
// are nearly contiguous) are copied into LDS once; each nonzero then carries a 16-bit
// position in that list (lidx) and its neighbour is an LDS read.  P is streamed as
// aligned 8-element chunks: 16 bytes of lidx and 32 bytes of fp32 values, non-temporal.
// A tile with more distinct neighbours than fit (ucap) reads its neighbours from HBM
// through the int32 columns instead.  Same arithmetic and update as attr_update_kernel;
// block_sums gets one record per tile.
// ---------------------------------------------------------------------------------
template <int D, int LANES, int BATCH, int NT>
__global__ void __launch_bounds__(NT)
attr_tile_kernel(int n, int ntiles, int tile_rows, const int *__restrict__ row_ptr,
                 const uint16_t *__restrict__ lidx, const int *__restrict__ col_idx,
                 const float *__restrict__ pval, const int *__restrict__ toff,
                 const int *__restrict__ tcols, int ucap,
                 const typename Vec<D>::T *__restrict__ Yg, const float *__restrict__ F,
                 double *__restrict__ Y64, double *__restrict__ uY, double *__restrict__ gains,
                 typename Vec<D>::T *__restrict__ Yg_next, const double *__restrict__ scal,
                 double exag4, double mom, double lr, double *__restrict__ block_sums)
{
    using V = typename Vec<D>::T;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    V *s_rec = reinterpret_cast<V *>(s_dyn);
    __shared__ double gsum[(NT / LANES)][3];
    const int sub = threadIdx.x & (LANES - 1);
    const int grp = threadIdx.x / LANES;
    // XCD-aware placement as in attr_update_kernel (gridDim.x is a multiple of 8;
    // tiles past ntiles are empty)
    const int lblk = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int row0 = min(n, lblk * tile_rows);
    const int row1 = min(n, row0 + tile_rows);
    const int u0 = (lblk < ntiles) ? toff[lblk] : 0;
    const int U = (lblk < ntiles) ? toff[lblk + 1] - u0 : 0;
    const bool local = U <= ucap;  // uniform over the workgroup
    if (local)
        for (int u = threadIdx.x; u < U; u += NT) s_rec[u] = Yg[tcols[u0 + u]];
    const double rep_scale = rep_scale_of(scal);
    const double m_d = (sub < D) ? scal[SLOT_MEAN + sub] : 0.0;
    double my_sum = 0.0;
    __syncthreads();

    for (int i = row0 + grp; i < row1; i += (NT / LANES)) {
        const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
        const size_t kd = (size_t)i * D + (sub < D ? sub : 0);
        double f_d = 0.0, u_d = 0.0, g_d = 0.0, y_d = 0.0;
        if (sub < D) {
            f_d = (double)F[kd];
            u_d = uY[kd];
            g_d = gains[kd];
            y_d = Y64[kd];
        }
        const V yi = Yg[i];
        float ax = 0.f, ay = 0.f, az = 0.f;
        auto accum = [&](const V &y, float p) {
            const float dx = yi.x - y.x, dy = yi.y - y.y;
            float r2 = fmaf(dx, dx, fmaf(dy, dy, 1.0f));
            float dz = 0.f;
            if constexpr (D == 3) { dz = yi.z - y.z; r2 = fmaf(dz, dz, r2); }
            const float q = p * fast_rcp(r2);
            ax = fmaf(q, dx, ax);
            ay = fmaf(q, dy, ay);
            if constexpr (D == 3) az = fmaf(q, dz, az);
        };
        if (local) {
            const int c1 = (e1 + 7) >> 3;  // 8-element chunks [e0/8, c1) cover the row
            for (int cb = (e0 >> 3) + sub; cb < c1; cb += BATCH * LANES) {
                v4i li[BATCH];
                v4f pa[BATCH], pb[BATCH];
#pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const int c = cb + t * LANES;
                    li[t] = v4i{0, 0, 0, 0};
                    pa[t] = pb[t] = v4f{0.f, 0.f, 0.f, 0.f};
                    if (c < c1) {
                        li[t] = __builtin_nontemporal_load((const v4i *)lidx + c);
                        pa[t] = __builtin_nontemporal_load((const v4f *)pval + 2 * c);
                        pb[t] = __builtin_nontemporal_load((const v4f *)pval + 2 * c + 1);
                    }
                }
#pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const int e = 8 * (cb + t * LANES);
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const bool ok = e + u >= e0 && e + u < e1;
                        const int l = (li[t][u >> 1] >> (16 * (u & 1))) & 0xffff;
                        const V y = s_rec[ok ? l : 0];
                        const float p = (u < 4) ? pa[t][u] : pb[t][u - 4];
                        accum(ok ? y : yi, ok ? p : 0.f);
                    }
                }
            }
        } else {
            const int c1 = (e1 + 3) >> 2;  // 4-element chunks of the int32 columns
            for (int cb = (e0 >> 2) + sub; cb < c1; cb += BATCH * LANES) {
                v4i cc[BATCH];
                v4f pv[BATCH];
#pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const int c = cb + t * LANES;
                    cc[t] = v4i{i, i, i, i};
                    pv[t] = v4f{0.f, 0.f, 0.f, 0.f};
                    if (c < c1) {
                        cc[t] = __builtin_nontemporal_load((const v4i *)col_idx + c);
                        pv[t] = __builtin_nontemporal_load((const v4f *)pval + c);
                    }
                }
#pragma unroll
                for (int t = 0; t < BATCH; t++) {
                    const int e = 4 * (cb + t * LANES);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const bool ok = e + u >= e0 && e + u < e1;
                        const V y = Yg[ok ? cc[t][u] : i];
                        accum(y, ok ? pv[t][u] : 0.f);
                    }
                }
            }
        }
        my_sum += row_epilogue<D, LANES>(ax, ay, az, sub, kd, f_d, u_d, g_d, y_d, m_d, exag4,
                                         rep_scale, mom, lr, Y64, uY, gains, Yg_next + i);
    }
    if (sub < D) gsum[grp][sub] = my_sum;
    __syncthreads();
    if (threadIdx.x < D) {
        double t = 0.0;
        for (int w = 0; w < (NT / LANES); w++) t += gsum[w][threadIdx.x];
        block_sums[(size_t)lblk * 3 + threadIdx.x] = t;
    }
}

// Update from a separately computed attractive sum A (the ATTR_ONLY kernel): the same
// gains/momentum rule, record and centring sums as the fused kernel, one thread per row,
// blocks over the same row ranges as the fused kernel.  The block's centring sum is added
// in the fused kernel's order (groups rows g, g + G, ... each summed in turn, then the
// groups in order, G = 256 / its lanes per row), so the split pass gives the fused pass's
// bits.
template <int D>
__global__ void __launch_bounds__(256)
update_kernel(int n, const float *__restrict__ A, const float *__restrict__ F,
              double *__restrict__ Y64, double *__restrict__ uY, double *__restrict__ gains,
              typename Vec<D>::T *__restrict__ Yg_next, const double *__restrict__ scal,
              double exag4, double mom, double lr, int rows_per_block,
              double *__restrict__ block_sums, int groups)
{
    __shared__ double red[3][256];
    const int lblk = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int row0 = lblk * rows_per_block;
    const int row1 = min(n, row0 + rows_per_block);
    const double rep_scale = rep_scale_of(scal);
    double m[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; d++) m[d] = scal[SLOT_MEAN + d];
    for (int i = row0 + threadIdx.x; i < row1; i += 256) {
        double yn[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < D; d++) {
            const size_t kd = (size_t)i * D + d;
            const double g = exag4 * (double)A[kd] - rep_scale * (double)F[kd];
            double u = uY[kd], gn = gains[kd];
            gn = (u * g < 0.0) ? gn + 0.2 : gn * 0.8;
            if (gn < 0.01) gn = 0.01;
            u = mom * u - lr * (gn * g);
            yn[d] = (Y64[kd] - m[d]) + u;
            Y64[kd] = yn[d];
            uY[kd] = u;
            gains[kd] = gn;
        }
        typename Vec<D>::T rec;
        rec.x = (float)yn[0];
        rec.y = (float)yn[1];
        if constexpr (D == 3) { rec.z = (float)yn[2]; rec.w = 0.f; }
        Yg_next[i] = rec;
    }
    __syncthreads();  // the block's new coordinates are written (read back below)
    if (threadIdx.x < groups) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int i = row0 + threadIdx.x; i < row1; i += groups)
            for (int d = 0; d < D; d++) s[d] += Y64[(size_t)i * D + d];
        for (int d = 0; d < 3; d++) red[d][threadIdx.x] = s[d];
    }
    __syncthreads();
    if (threadIdx.x < D) {
        double t = 0.0;
        for (int w = 0; w < groups; w++) t += red[threadIdx.x][w];
        block_sums[(size_t)lblk * 3 + threadIdx.x] = t;
    }
}

// Sum of the per-block sums -> new pending mean (SLOT_MEAN..+D-1), in two launches with a
// fixed order of additions: kMeanParts workgroups each add a contiguous run of the block
// records (thread t the records t, t + 256, ... of the run, then a fixed tree) into
// partials stored after the records; one workgroup adds the partials (fixed tree) and
// folds a finiteness flag (SLOT_BAD) for the host.  The block-sum buffer holds
// kMeanParts records beyond the blocks of any launch.
static_assert(kMeanParts == 256, "one partial per thread of the final workgroup");

__device__ __forceinline__ void tree_sum3_256(double (&a)[3], double (*part)[256], int D)
{
    const int t = threadIdx.x;
#pragma unroll
    for (int d = 0; d < 3; d++) part[d][t] = a[d];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
#pragma unroll
            for (int d = 0; d < 3; d++)
                if (d < D) part[d][t] += part[d][t + off];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
mean_partial_kernel(int nblocks, int D, double *__restrict__ block_sums)
{
    __shared__ double part[3][256];
    const int per = (nblocks + kMeanParts - 1) / kMeanParts;
    const int b0 = blockIdx.x * per, b1 = min(nblocks, b0 + per);
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = b0 + (int)threadIdx.x; b < b1; b += 256)
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (d < D) a[d] += block_sums[(size_t)b * 3 + d];
    tree_sum3_256(a, part, D);
    if (threadIdx.x < 3) block_sums[((size_t)nblocks + blockIdx.x) * 3 + threadIdx.x] = part[threadIdx.x][0];
}

__global__ void __launch_bounds__(256)
mean_finalize_kernel(int nblocks, int D, int n, const double *__restrict__ block_sums,
                     double *__restrict__ scal)
{
    __shared__ double part[3][256];
    const int t = threadIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (d < D) a[d] = block_sums[((size_t)nblocks + t) * 3 + d];  // kMeanParts == 256
    tree_sum3_256(a, part, D);
    if (t < D) {
        const double m = part[t][0] / (double)n;
        scal[SLOT_MEAN + t] = m;
        if (!isfinite(m)) scal[SLOT_BAD] = 1.0;
    }
}

int launch_mean(int D, int n, int nblocks, double *block_sums, double *scal, hipStream_t st)
{
    hipLaunchKernelGGL(mean_partial_kernel, dim3(kMeanParts), dim3(256), 0, st, nblocks, D, block_sums);
    hipLaunchKernelGGL(mean_finalize_kernel, dim3(1), dim3(256), 0, st, nblocks, D, n, block_sums, scal);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Direct O(N^2) repulsive sums in FP32 (the reference's direct branch,
// src/tsne_h2pack.c:407-491, used below 2000 points and for method='exact').
// One thread per target; sources staged through LDS in tiles of 256.  Produces
//   s_i = sum_{j != i} K_ij,   F_i = sum_{j != i} K_ij (y_i - y_j),  K = (1+r^2)^-2.
// F is formed from differences directly, which avoids the y*s - v cancellation of
// the reference's assembly (src/tsne_h2pack.c:619) in float32.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
repulsive_direct_f32(int n, const typename Vec<D>::T *__restrict__ Yg, float *__restrict__ s,
                     float *__restrict__ F)
{
    using V = typename Vec<D>::T;
    __shared__ V tile[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    V yi;
    if (i < n) yi = Yg[i]; else { yi.x = yi.y = 0.f; if constexpr (D == 3) { yi.z = yi.w = 0.f; } }
    float si = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
    for (int t0 = 0; t0 < n; t0 += 256) {
        __syncthreads();
        if (t0 + threadIdx.x < n) tile[threadIdx.x] = Yg[t0 + threadIdx.x];
        __syncthreads();
        const int tn = min(256, n - t0);
        for (int u = 0; u < tn; u++) {
            const V yj = tile[u];
            const float dx = yi.x - yj.x, dy = yi.y - yj.y;
            float r2 = fmaf(dx, dx, fmaf(dy, dy, 1.0f));
            float dz = 0.f;
            if constexpr (D == 3) { dz = yi.z - yj.z; r2 = fmaf(dz, dz, r2); }
            float w = fast_rcp(r2);
            w = (t0 + u == i) ? 0.f : w * w;
            si += w;
            fx = fmaf(w, dx, fx);
            fy = fmaf(w, dy, fy);
            if constexpr (D == 3) fz = fmaf(w, dz, fz);
        }
    }
    if (i < n) {
        s[i] = si;
        F[(size_t)i * D] = fx;
        F[(size_t)i * D + 1] = fy;
        if constexpr (D == 3) F[(size_t)i * D + 2] = fz;
    }
}

// ---------------------------------------------------------------------------------
// Z = sum_i s_i + 2 sum_i (y_i - m) . F_i  (FP64 accumulation, fixed order).
// This is the reference's identity Z = sum s + 2 sum |y|^2 s - 2 sum y.v
// (src/tsne_h2pack.c:554-590) rewritten with F_i = y_i s_i - v_i, which keeps the two
// large terms from cancelling.  Points are read through Yg (the coordinates the
// product used) with the pending mean m removed.  Stage 1: per-block partials.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
z_partial_kernel(int n, const typename Vec<D>::T *__restrict__ Yg, const float *__restrict__ s,
                 const float *__restrict__ F, const double *__restrict__ scal,
                 int per_block, double *__restrict__ partial)
{
    __shared__ double red[256];
    const int b0 = blockIdx.x * per_block, b1 = min(n, b0 + per_block);
    double acc = 0.0;
    const double m0 = scal[SLOT_YSHIFT], m1 = scal[SLOT_YSHIFT + 1];
    const double m2 = (D == 3) ? scal[SLOT_YSHIFT + 2] : 0.0;
    for (int i = b0 + threadIdx.x; i < b1; i += 256) {
        const auto y = Yg[i];
        double yf = ((double)y.x - m0) * (double)F[(size_t)i * D] +
                    ((double)y.y - m1) * (double)F[(size_t)i * D + 1];
        if constexpr (D == 3) yf += ((double)y.z - m2) * (double)F[(size_t)i * D + 2];
        acc += (double)s[i] + 2.0 * yf;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Stage 2: partials in order -> Z (256 threads take contiguous chunks, then the chunk
// sums are added in order).  A non-positive or non-finite Z is replaced by 1 and
// flagged, as the reference returns 1.0 and continues (src/tsne_h2pack.c:604-611).
__global__ void __launch_bounds__(256)
z_finalize_kernel(int nblocks, const double *__restrict__ partial, double *__restrict__ scal)
{
    __shared__ double part[256];
    const int t = threadIdx.x;
    const int per = (nblocks + 255) / 256;
    double a = 0.0;
    for (int b = t * per; b < min(nblocks, (t + 1) * per); b++) a += partial[b];
    part[t] = a;
    __syncthreads();
    if (t == 0) {
        double z = 0.0;
        for (int k = 0; k < 256; k++) z += part[k];
        const bool bad = !(z > 0.0) || !isfinite(z);
        if (bad) { scal[SLOT_ZBAD] = 1.0; z = 1.0; }
        scal[SLOT_Z] = z;
        scal[SLOT_ZSKIP] = bad ? 1.0 : 0.0;
    }
}

// Copy the pending mean into the slot the Z pass subtracts (taken when the
// repulsive product was formed).
__global__ void shift_snapshot_kernel(double *__restrict__ scal)
{
    if (threadIdx.x < 3) scal[SLOT_YSHIFT + threadIdx.x] = scal[SLOT_MEAN + threadIdx.x];
}

// ---------------------------------------------------------------------------------
// KL divergence, src/tsne_h2pack.c:1870-1895: per row sum of p*exag*log(p*exag/q) with
// q = 1/((1+d^2) Z) over terms with p*exag > 1e-12 and q > 1e-12, in FP64; one
// wavefront per row; rows summed in fixed blocks.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
kl_rows_kernel(int n, const int *__restrict__ row_ptr, const int *__restrict__ col_idx,
               const double *__restrict__ pval64, const double *__restrict__ Y64,
               const double *__restrict__ scal, double exag, double *__restrict__ row_kl)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double Z = scal[SLOT_Z];
    double yi[3];
    for (int d = 0; d < D; d++) yi[d] = Y64[(size_t)i * D + d];
    double acc = 0.0;
    for (int e = row_ptr[i] + lane; e < row_ptr[i + 1]; e += 64) {
        const int j = col_idx[e];
        double r2 = 0.0;
        for (int d = 0; d < D; d++) {
            const double df = yi[d] - Y64[(size_t)j * D + d];
            r2 += df * df;
        }
        const double p = pval64[e] * exag;
        const double q = 1.0 / ((1.0 + r2) * Z);
        if (p > 1e-12 && q > 1e-12) acc += p * log(p / q);
    }
    acc = wave_sum(acc);
    if (lane == 0) row_kl[i] = acc;
}

__global__ void __launch_bounds__(256)
sum_blocks_kernel(int n, const double *__restrict__ x, int per_block, double *__restrict__ partial)
{
    __shared__ double red[256];
    const int b0 = blockIdx.x * per_block, b1 = min(n, b0 + per_block);
    double acc = 0.0;
    for (int i = b0 + threadIdx.x; i < b1; i += 256) acc += x[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
sum_finalize_kernel(int nblocks, const double *__restrict__ partial, double *__restrict__ dst)
{
    __shared__ double part[256];
    const int t = threadIdx.x;
    const int per = (nblocks + 255) / 256;
    double a = 0.0;
    for (int b = t * per; b < min(nblocks, (t + 1) * per); b++) a += partial[b];
    part[t] = a;
    __syncthreads();
    if (t == 0) {
        double z = 0.0;
        for (int k = 0; k < 256; k++) z += part[k];
        *dst = z;
    }
}

// ---------------------------------------------------------------------------------
// FP64 all-pairs check path: s_i and F_i = sum_{j != i} K_ij (y_i - y_j) in FP64 from
// the FP64 master coordinates (pending mean is a common shift and cancels).  Used to
// measure the error of the product path, not inside the optimisation loop.
// ---------------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
repulsive_direct_f64(int n, const double *__restrict__ Y64, const int *__restrict__ rows,
                     int n_rows, double *__restrict__ s, double *__restrict__ F)
{
    __shared__ double tile[256][D];
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int i = (r < n_rows) ? (rows ? rows[r] : r) : -1;
    double yi[3] = {0.0, 0.0, 0.0};
    if (i >= 0) for (int d = 0; d < D; d++) yi[d] = Y64[(size_t)i * D + d];
    double si = 0.0, f[3] = {0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < n; t0 += 256) {
        __syncthreads();
        if (t0 + threadIdx.x < n)
            for (int d = 0; d < D; d++) tile[threadIdx.x][d] = Y64[(size_t)(t0 + threadIdx.x) * D + d];
        __syncthreads();
        const int tn = min(256, n - t0);
        for (int u = 0; u < tn; u++) {
            double df[3], r2 = 1.0;
            for (int d = 0; d < D; d++) { df[d] = yi[d] - tile[u][d]; r2 += df[d] * df[d]; }
            double w = 1.0 / r2;
            w = (t0 + u == i) ? 0.0 : w * w;
            si += w;
            for (int d = 0; d < D; d++) f[d] += w * df[d];
        }
    }
    if (i >= 0) {
        s[r] = si;
        for (int d = 0; d < D; d++) F[(size_t)r * D + d] = f[d];
    }
}

// Z in FP64 from FP64 sums: Z = sum s + 2 sum (y - mean(y)) . F  (all rows).
template <int D>
__global__ void __launch_bounds__(256)
z_partial_f64_kernel(int n, const double *__restrict__ Y64, const double *__restrict__ s,
                     const double *__restrict__ F, const double *__restrict__ scal,
                     int per_block, double *__restrict__ partial)
{
    __shared__ double red[256];
    const int b0 = blockIdx.x * per_block, b1 = min(n, b0 + per_block);
    double acc = 0.0;
    for (int i = b0 + threadIdx.x; i < b1; i += 256) {
        double yf = 0.0;
        for (int d = 0; d < D; d++)
            yf += (Y64[(size_t)i * D + d] - scal[SLOT_MEAN + d]) * F[(size_t)i * D + d];
        acc += s[i] + 2.0 * yf;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------------
// Perplexity bisection, src/tsne_utils.c:952-1105, FP64, one wavefront per row.
// Distances are squared and shifted by the row minimum; beta starts at 1 and is
// doubled / halved until bracketed, then bisected, until |H - ln(perp)| < 1e-5 nats
// or 200 steps; P = exp(-beta (d^2 - dmin)) / sum in the row's storage order.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
perplexity_kernel(int n, const double *__restrict__ dist, const int *__restrict__ indptr,
                  double perplexity, double *__restrict__ p_out)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int b = indptr[i], k = indptr[i + 1] - b;
    if (k <= 0) return;
    double dmin = INFINITY;
    for (int j = lane; j < k; j += 64) { const double d = dist[b + j]; dmin = fmin(dmin, d * d); }
    for (int off = 32; off > 0; off >>= 1) dmin = fmin(dmin, __shfl_xor(dmin, off, 64));
    const double target = log(perplexity);
    double beta = 1.0, lo = 0.0, hi = 0.0;
    bool have_lo = false, have_hi = false;
    double sp = 1.0;
    for (int step = 0; step < 200; step++) {
        double a = 0.0, c = 0.0;
        for (int j = lane; j < k; j += 64) {
            const double d0 = dist[b + j];
            const double dd = d0 * d0 - dmin;
            const double pj = exp(-beta * dd);
            a += pj;
            c += dd * pj;
        }
        sp = wave_sum(a);
        const double sdp = wave_sum(c);
        const double H = log(sp) + beta * sdp / sp;
        const double diff = H - target;
        if (fabs(diff) < 1e-5) break;
        if (step == 199) break;
        if (diff > 0.0) {
            lo = beta; have_lo = true;
            beta = have_hi ? 0.5 * (beta + hi) : 2.0 * beta;
        } else {
            hi = beta; have_hi = true;
            beta = have_lo ? 0.5 * (beta + lo) : 0.5 * beta;
        }
    }
    const double inv = 1.0 / sp;
    for (int j = lane; j < k; j += 64) {
        const double d0 = dist[b + j];
        p_out[b + j] = exp(-beta * (d0 * d0 - dmin)) * inv;
    }
}

// ---------------------------------------------------------------------------------
// Host-side launchers (declared in kernels.h).
// ---------------------------------------------------------------------------------
namespace {
template <int D, int L, int BT, bool VEC>
void attr_launch(int nblocks, hipStream_t st, int n, const int *row_ptr, const int *col_idx,
                 const float *pval, const void *Yg, const float *F, double *Y64, double *uY,
                 double *gains, void *Yg_next, const double *scal, double exag, double mom,
                 double lr, int rpb, double *block_sums)
{
    using V = typename Vec<D>::T;
    hipLaunchKernelGGL((attr_update_kernel<D, L, BT, VEC>), dim3(nblocks), dim3(256), 0, st, n,
                       row_ptr, col_idx, pval, (const V *)Yg, F, Y64, uY, gains, (V *)Yg_next,
                       scal, 4.0 * exag, mom, lr, rpb, block_sums);
}
}  // namespace

// Row width of the attractive kernel: 8, 16 (default) or 32 lanes per row
// (environment TSNE_ATTR_LANES, read once).
static int attr_lanes()
{
    static int lanes = [] {
        const char *e = getenv("TSNE_ATTR_LANES");
        const int v = e ? atoi(e) : 16;
        return (v == 8 || v == 32) ? v : 16;
    }();
    return lanes;
}

// Chunked non-temporal reads of P (default) or per-element reads (TSNE_ATTR_VEC=0, which
// also honours TSNE_ATTR_LANES); read once.
static bool attr_vec()
{
    static bool v = [] {
        const char *e = getenv("TSNE_ATTR_VEC");
        return !(e && e[0] == '0');
    }();
    return v;
}

int launch_attr_tile(int D, int n, const AttrTiles &tl, int nt, const int *row_ptr, const int *col_idx,
                     const float *pval, const void *Yg, const float *F, double *Y64, double *uY,
                     double *gains, void *Yg_next, const double *scal, double exag, double mom,
                     double lr, double *block_sums, int nblocks, hipStream_t st,
                     hipEvent_t ev_begin, hipEvent_t ev_end)
{
    const int grid = (tl.ntiles + 7) & ~7;
    if (grid > nblocks) return -1;  // block_sums holds nblocks records
    const int ucap = tl.umax;       // tiles with more distinct neighbours gather from HBM
    const size_t lds = (size_t)std::max(ucap, 1) * (D == 2 ? sizeof(float2) : sizeof(float4));
    if (lds > (size_t)kTileLdsMaxKB * 1024) return -1;  // beyond the CU's LDS with the static part
    {
        // the device's per-workgroup LDS limit (read once): a launch beyond it would be
        // rejected by the runtime, so refuse it here
        static int lim = [] {
            int dev = 0, v = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
                v = 65536;
            return v;
        }();
        if (lds + 4096 > (size_t)lim) {
            fprintf(stderr, "tsne_h2pack: tile LDS %zu B exceeds the %d B per workgroup\n", lds, lim);
            return -1;
        }
    }
    if (ev_begin) (void)hipEventRecord(ev_begin, st);
#define TSNE_TILE_ARGS dim3(grid), dim3(NT_), lds, st, n, tl.ntiles, tl.rows, row_ptr, tl.lidx, \
        col_idx, pval, tl.off, tl.cols, ucap
    if (D == 2) {
        if (nt == 1024) {
            constexpr int NT_ = 1024;
            if (lds > 65536)
                (void)hipFuncSetAttribute((const void *)attr_tile_kernel<2, 16, 2, 1024>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((attr_tile_kernel<2, 16, 2, 1024>), TSNE_TILE_ARGS, (const float2 *)Yg,
                               F, Y64, uY, gains, (float2 *)Yg_next, scal, 4.0 * exag, mom, lr,
                               block_sums);
        } else {
            constexpr int NT_ = 256;
            if (lds > 65536)
                (void)hipFuncSetAttribute((const void *)attr_tile_kernel<2, 16, 2, 256>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((attr_tile_kernel<2, 16, 2, 256>), TSNE_TILE_ARGS, (const float2 *)Yg,
                               F, Y64, uY, gains, (float2 *)Yg_next, scal, 4.0 * exag, mom, lr,
                               block_sums);
        }
    } else {
        if (nt == 1024) {
            constexpr int NT_ = 1024;
            if (lds > 65536)
                (void)hipFuncSetAttribute((const void *)attr_tile_kernel<3, 16, 2, 1024>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((attr_tile_kernel<3, 16, 2, 1024>), TSNE_TILE_ARGS, (const float4 *)Yg,
                               F, Y64, uY, gains, (float4 *)Yg_next, scal, 4.0 * exag, mom, lr,
                               block_sums);
        } else {
            constexpr int NT_ = 256;
            if (lds > 65536)
                (void)hipFuncSetAttribute((const void *)attr_tile_kernel<3, 16, 2, 256>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((attr_tile_kernel<3, 16, 2, 256>), TSNE_TILE_ARGS, (const float4 *)Yg,
                               F, Y64, uY, gains, (float4 *)Yg_next, scal, 4.0 * exag, mom, lr,
                               block_sums);
        }
    }
#undef TSNE_TILE_ARGS
    if (ev_end) (void)hipEventRecord(ev_end, st);
    if (launch_mean(D, n, grid, block_sums, (double *)scal, st) != 0) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_attr_update(int D, int n, const int *row_ptr, const int *col_idx, const float *pval,
                       const void *Yg, const float *F, double *Y64, double *uY, double *gains,
                       void *Yg_next, const double *scal, double exag, double mom, double lr,
                       double *block_sums, int nblocks, hipStream_t st, hipEvent_t ev_begin,
                       hipEvent_t ev_end)
{
    const int rpb = (n + nblocks - 1) / nblocks;
    const int L = attr_lanes();
    if (ev_begin) (void)hipEventRecord(ev_begin, st);
#define TSNE_ATTR_ARGS nblocks, st, n, row_ptr, col_idx, pval, Yg, F, Y64, uY, gains, Yg_next, \
                       scal, exag, mom, lr, rpb, block_sums
    if (attr_vec()) {
        if (D == 2) attr_launch<2, 16, 3, true>(TSNE_ATTR_ARGS);
        else attr_launch<3, 16, 3, true>(TSNE_ATTR_ARGS);
    } else if (D == 2) {
        if (L == 8) attr_launch<2, 8, 12, false>(TSNE_ATTR_ARGS);
        else if (L == 32) attr_launch<2, 32, 6, false>(TSNE_ATTR_ARGS);
        else attr_launch<2, 16, 10, false>(TSNE_ATTR_ARGS);
    } else {
        if (L == 8) attr_launch<3, 8, 12, false>(TSNE_ATTR_ARGS);
        else if (L == 32) attr_launch<3, 32, 6, false>(TSNE_ATTR_ARGS);
        else attr_launch<3, 16, 10, false>(TSNE_ATTR_ARGS);
    }
#undef TSNE_ATTR_ARGS
    if (ev_end) (void)hipEventRecord(ev_end, st);
    if (launch_mean(D, n, nblocks, block_sums, (double *)scal, st) != 0) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_attr_only(int D, int n, const int *row_ptr, const int *col_idx, const float *pval,
                     const void *Yg, float *A, int nblocks, int grid, hipStream_t st)
{
    const int rpb = (n + nblocks - 1) / nblocks;
    // grid: workgroups of the launch (0: one per logical block); a multiple of 8, so a
    // workgroup's logical blocks all sit on its XCD
    grid = (grid <= 0 || grid >= nblocks) ? nblocks : std::max(8, grid & ~7);
    if (D == 2)
        hipLaunchKernelGGL((attr_update_kernel<2, 16, 3, true, true>), dim3(grid), dim3(256), 0, st,
                           n, row_ptr, col_idx, pval, (const float2 *)Yg, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, 0.0, 0.0, 0.0, rpb, nullptr, A, nblocks);
    else
        hipLaunchKernelGGL((attr_update_kernel<3, 16, 3, true, true>), dim3(grid), dim3(256), 0, st,
                           n, row_ptr, col_idx, pval, (const float4 *)Yg, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, 0.0, 0.0, 0.0, rpb, nullptr, A, nblocks);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_update(int D, int n, const float *A, const float *F, double *Y64, double *uY,
                  double *gains, void *Yg_next, const double *scal, double exag, double mom,
                  double lr, double *block_sums, int nblocks, hipStream_t st)
{
    const int rpb = (n + nblocks - 1) / nblocks;
    // row groups of the fused kernel (its centring-sum order): 16 lanes per row in the
    // chunked form, else the configured row width
    const int groups = 256 / (attr_vec() ? 16 : attr_lanes());
    if (D == 2)
        hipLaunchKernelGGL(update_kernel<2>, dim3(nblocks), dim3(256), 0, st, n, A, F, Y64, uY,
                           gains, (float2 *)Yg_next, scal, 4.0 * exag, mom, lr, rpb, block_sums,
                           groups);
    else
        hipLaunchKernelGGL(update_kernel<3>, dim3(nblocks), dim3(256), 0, st, n, A, F, Y64, uY,
                           gains, (float4 *)Yg_next, scal, 4.0 * exag, mom, lr, rpb, block_sums,
                           groups);
    if (launch_mean(D, n, nblocks, block_sums, (double *)scal, st) != 0) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_repulsive_direct_f32(int D, int n, const void *Yg, float *s, float *F, hipStream_t st)
{
    const int g = (n + 255) / 256;
    if (D == 2)
        hipLaunchKernelGGL(repulsive_direct_f32<2>, dim3(g), dim3(256), 0, st, n,
                           (const float2 *)Yg, s, F);
    else
        hipLaunchKernelGGL(repulsive_direct_f32<3>, dim3(g), dim3(256), 0, st, n,
                           (const float4 *)Yg, s, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_z(int D, int n, const void *Yg, const float *s, const float *F, double *scal,
             double *partial, int nblocks, hipStream_t st)
{
    const int per = (n + nblocks - 1) / nblocks;
    if (D == 2)
        hipLaunchKernelGGL(z_partial_kernel<2>, dim3(nblocks), dim3(256), 0, st, n,
                           (const float2 *)Yg, s, F, scal, per, partial);
    else
        hipLaunchKernelGGL(z_partial_kernel<3>, dim3(nblocks), dim3(256), 0, st, n,
                           (const float4 *)Yg, s, F, scal, per, partial);
    hipLaunchKernelGGL(z_finalize_kernel, dim3(1), dim3(256), 0, st, nblocks, partial, scal);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// P's FP32 values from its FP64 values (round to nearest, as a host cast), and a constant
// fill: both on the device, so the workspace set-up uploads P once.
__global__ void __launch_bounds__(256)
f64_to_f32_kernel(const double *__restrict__ src, float *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = (float)src[i];
}

__global__ void __launch_bounds__(256) fill_f64_kernel(double *__restrict__ dst, double v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = v;
}

int launch_f64_to_f32(const double *src, float *dst, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    const int g = (int)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(g), dim3(256), 0, st, src, dst, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fill_f64(double *dst, double v, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    const int g = (int)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL(fill_f64_kernel, dim3(g), dim3(256), 0, st, dst, v, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_shift_snapshot(double *scal, hipStream_t st)
{
    hipLaunchKernelGGL(shift_snapshot_kernel, dim3(1), dim3(64), 0, st, scal);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_kl(int D, int n, const int *row_ptr, const int *col_idx, const double *pval64,
              const double *Y64, const double *scal, double exag, double *row_kl,
              double *partial, int nblocks, double *dst, hipStream_t st)
{
    const int g = (n + 3) / 4;
    if (D == 2)
        hipLaunchKernelGGL(kl_rows_kernel<2>, dim3(g), dim3(256), 0, st, n, row_ptr, col_idx,
                           pval64, Y64, scal, exag, row_kl);
    else
        hipLaunchKernelGGL(kl_rows_kernel<3>, dim3(g), dim3(256), 0, st, n, row_ptr, col_idx,
                           pval64, Y64, scal, exag, row_kl);
    const int per = (n + nblocks - 1) / nblocks;
    hipLaunchKernelGGL(sum_blocks_kernel, dim3(nblocks), dim3(256), 0, st, n, row_kl, per, partial);
    hipLaunchKernelGGL(sum_finalize_kernel, dim3(1), dim3(256), 0, st, nblocks, partial, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_repulsive_direct_f64(int D, int n, const double *Y64, const int *rows, int n_rows,
                                double *s, double *F, hipStream_t st)
{
    const int g = (n_rows + 255) / 256;
    if (D == 2)
        hipLaunchKernelGGL(repulsive_direct_f64<2>, dim3(g), dim3(256), 0, st, n, Y64, rows,
                           n_rows, s, F);
    else
        hipLaunchKernelGGL(repulsive_direct_f64<3>, dim3(g), dim3(256), 0, st, n, Y64, rows,
                           n_rows, s, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_z_f64(int D, int n, const double *Y64, const double *s, const double *F,
                 const double *scal, double *partial, int nblocks, double *dst, hipStream_t st)
{
    const int per = (n + nblocks - 1) / nblocks;
    if (D == 2)
        hipLaunchKernelGGL(z_partial_f64_kernel<2>, dim3(nblocks), dim3(256), 0, st, n, Y64, s,
                           F, scal, per, partial);
    else
        hipLaunchKernelGGL(z_partial_f64_kernel<3>, dim3(nblocks), dim3(256), 0, st, n, Y64, s,
                           F, scal, per, partial);
    hipLaunchKernelGGL(sum_finalize_kernel, dim3(1), dim3(256), 0, st, nblocks, partial, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_perplexity(int n, const double *dist, const int *indptr, double perplexity,
                      double *p_out, hipStream_t st)
{
    const int g = (n + 3) / 4;
    hipLaunchKernelGGL(perplexity_kernel, dim3(g), dim3(256), 0, st, n, dist, indptr, perplexity,
                       p_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsne
