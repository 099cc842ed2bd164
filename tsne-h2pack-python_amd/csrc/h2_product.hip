// h2_product.hip -- the H2 product [s | F] of the repulsive kernel, after the upward
// pass (fused into the ID kernel):
//   coupling   z_i  = sum_ADM K(J_i, J_j) [w_j0, (c_i - c_j) w_j0 - w_jY]
//                   + sum_X   K(J_i, y_u) [1,    c_i - y_u]
//   downward   z_c += U_p[rows of c] z_p, with the F part moved by (c_c - c_p) t
//   leaves     s_t  = (U_i z_i0)_t + near + W,  F_t = (y_t - c_i)(U_i z_i0)_t + (U_i z_iY)_t + ...
// Coefficients are kept relative to node centres, so every term is formed from
// differences (y_t - c_i, c_i - c_j, y_t - y_u) and no large-coordinate cancellation
// occurs; F_t = sum_u K(y_t, y_u)(y_t - y_u) directly.
//
// Reference counterpart: the coupling / near / downward sweeps of the fused product
// tsne_h2_matmul (src/tsne_h2_matmul.c:336-440) and the assembly of s, v
// (src/tsne_h2pack.c:538-546), which this formulation replaces.
#include "h2_internal.h"

namespace tsne {

namespace {

template <int D> using VT = typename Vec<D>::T;

__device__ __forceinline__ float kern32(float dx, float dy, float dz)
{
    const float w = fast_rcp(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, 1.0f))));
    return w * w;
}

__device__ __forceinline__ int rank_stride(int r)
{
    return r <= 8 ? 8 : (r <= 16 ? 16 : (r <= 32 ? 32 : 64));
}

__device__ __forceinline__ float4 group_sum(float4 v, int RS)
{
    for (int off = RS; off < 64; off <<= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
        v.z += __shfl_xor(v.z, off, 64);
        v.w += __shfl_xor(v.w, off, 64);
    }
    return v;
}

}  // namespace

// Sources of the coupling of target node i: the skeleton points of its ADM nodes, then
// the points of its X leaves.  Source s carries its coordinates and its row of W':
// [w0, (c_i - c_j) w0 - wY] for skeleton point q of ADM node j, [1, c_i - y] for point y
// of an X leaf.  A wavefront walks the members in rounds of up to 64 (one per lane; their
// source counts scanned into off[0..64]) and stages 64 sources at a time into LDS.
template <int D>
__device__ __forceinline__ int member_round(const H2Dev &h, int m0, int na, int nm,
                                            int a0, int x0, int lane, int *off, int *mj)
{
    const int m = m0 + lane;
    int sz = 0, j = -1;
    if (m < nm) {
        j = (m < na) ? (int)(uint32_t)h.adm[a0 + m] : (int)(uint32_t)h.xl[x0 + m - na];
        sz = (m < na) ? h.rank[j] : (h.ne[j] - h.nb[j]);
    }
    int x = sz;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    off[lane + 1] = x;
    if (lane == 0) off[0] = 0;
    mj[lane] = j;
    __builtin_amdgcn_wave_barrier();
    return __shfl(x, 63, 64);
}

template <int D>
__device__ __forceinline__ void stage_source(const H2Dev &h, float4 ci, int m0, int na,
                                             int nmem, int sidx, int T, const int *off,
                                             const int *mj, float4 &xs, float4 &ws)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    xs = ci;
    ws = make_float4(0.f, 0.f, 0.f, 0.f);
    if (sidx >= T) return;
    int lo = 0, hi = nmem;  // member a with off[a] <= sidx < off[a+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= sidx) lo = mid; else hi = mid;
    }
    const int jj = mj[lo], q = sidx - off[lo];
    if (m0 + lo < na) {
        const float4 cj = h.cent[jj];
        xs = h.skelf[(size_t)jj * RMAX + q];
        const float4 w = h.wc[(size_t)jj * RMAX + q];
        ws = make_float4(w.x, (ci.x - cj.x) * w.x - w.y, (ci.y - cj.y) * w.x - w.z,
                         (ci.z - cj.z) * w.x - w.w);
    } else {
        const VT<D> y = ((const VT<D> *)h.Ys)[h.nb[jj] + q];
        xs.x = y.x;
        xs.y = y.y;
        xs.z = 0.f;
        if constexpr (D == 3) xs.z = y.z;
        ws = make_float4(1.f, ci.x - y.x, ci.y - y.y, (D == 3) ? ci.z - xs.z : 0.f);
    }
}

// FP32 coupling of target node i by one wavefront; lanes are (group g, row k): row k of
// the node's skeleton, the groups split each staged round of 64 sources, which every lane
// reads from LDS (the wavefront's own staging arrays).
template <int D>
__device__ __forceinline__ void couple_node_f32(const H2Dev &h, int i, int r, int lane,
                                                float4 *s_x, float4 *s_w, int *s_off, int *s_j)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    constexpr int CH = 64;
    const int RS = rank_stride(r), G = 64 / RS;
    const int k = lane % RS, g = lane / RS;
    const float4 ci = h.cent[i];
    const float4 xk = (k < r) ? h.skelf[(size_t)i * RMAX + k] : ci;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int a0 = h.adm_beg[i], a1 = h.adm_end[i], x0 = h.x_beg[i], x1 = h.x_end[i];
    const int na = a1 - a0, nm = na + (x1 - x0);
    for (int m0 = 0; m0 < nm; m0 += 64) {
        const int T = member_round<D>(h, m0, na, nm, a0, x0, lane, s_off, s_j);
        const int nmem = min(64, nm - m0);
        for (int c0 = 0; c0 < T; c0 += CH) {
            float4 xs, ws;
            stage_source<D>(h, ci, m0, na, nmem, c0 + lane, T, s_off, s_j, xs, ws);
            s_x[lane] = xs;
            s_w[lane] = ws;
            __builtin_amdgcn_wave_barrier();
            const int cn = min(CH, T - c0);
            for (int sl = g; sl < cn; sl += G) {
                const float4 q = s_x[sl];
                const float4 w = s_w[sl];
                const float K = kern32(xk.x - q.x, xk.y - q.y, D == 3 ? xk.z - q.z : 0.f);
                acc.x = fmaf(K, w.x, acc.x);
                acc.y = fmaf(K, w.y, acc.y);
                acc.z = fmaf(K, w.z, acc.z);
                if constexpr (D == 3) acc.w = fmaf(K, w.w, acc.w);
            }
            __builtin_amdgcn_wave_barrier();  // the next round overwrites the staging
        }
        __builtin_amdgcn_wave_barrier();
    }
    acc = group_sum(acc, RS);
    if (g == 0 && k < r) h.zc[(size_t)i * RMAX + k] = acc;
}

// FP32 coupling: one wavefront per target node (all non-root nodes).
template <int D>
__global__ void __launch_bounds__(256)
couple_kernel(H2Dev h, int n_nodes)
{
    __shared__ float4 s_x[4][64], s_w[4][64];
    __shared__ int s_off[4][65], s_j[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = 1 + blockIdx.x * 4 + wv;
    if (i >= n_nodes) return;  // per wavefront: only wavefront-level LDS traffic below
    const int r = h.rank[i];
    if (r <= 0) return;
    couple_node_f32<D>(h, i, r, lane, s_x[wv], s_w[wv], s_off[wv], s_j[wv]);
}

// ---------------------------------------------------------------------------------
// Mixed-precision coupling: the same z_i as couple_kernel, as a batched small GEMM on
// MFMA.  For target node i (one wavefront) the sources are the skeleton points of its
// ADM nodes and the points of its X leaves, concatenated; with W'_s the source's row
// [w0, (c_i - c_j) w0 - wY] (ADM) or [1, c_i - y] (X),
//     z_i = K(J_i, sources) W',
// an (r_i x S) by (S x 4) product.  v_mfma_f32_4x4x4bf16_1k computes 16 independent
// 4x4x4 blocks: block b = lane / 4 takes target rows 4g..4g+3 (g = b mod G, G = the
// rank in groups of 4, rounded up to a power of two) against 4 sources of phase
// b / G, so one instruction consumes 64 / G sources; the phases are summed at the end.
// Lane u of a block supplies row 4g+u of K (four kernel values, computed in FP32 and
// rounded to BF16) and column u of W' for the block's four sources, split into two
// BF16 terms (hi + lo) so that W' keeps ~16 bits; accumulation is FP32.
//
// Error control: rounding K to BF16 perturbs each term by at most 2^-9 relative with
// signs that vary over the sources, so the error of an entry of z is estimated as
// 2^-9 sqrt(sum (K W')^2) (a third MFMA accumulates K^2 W'^2).  A node where the estimate
// exceeds tau times the entry's scale for any entry (strong cancellation between its
// sources) is flagged and redone at once by the same wavefront in FP32 (couple_node_f32);
// the scale is |z| for the s column and max(|z|, L |z_0|) for a force column (L: the
// node's half-side).
// ---------------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// four floats to BF16 (round to nearest even; two v_cvt_pk_bf16_f32)
__device__ __forceinline__ v4s bf16x4(v4f x)
{
    return __builtin_bit_cast(v4s, __builtin_convertvector(x, v4bf));
}

__device__ __forceinline__ v4f f32x4(v4s x)
{
    return __builtin_convertvector(__builtin_bit_cast(v4bf, x), v4f);
}

template <int D>
__global__ void __launch_bounds__(256)
couple_mfma_kernel(H2Dev h, int n_nodes, float tau)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    constexpr int CH = 64;  // sources staged per round
    __shared__ float4 s_x[4][CH];
    // W' of the staged sources as BF16, column-major so that a lane's four sources of
    // one column are one 8-byte read: hi and lo terms, and W'^2 for the error estimate
    __shared__ __align__(8) short s_wh[4][4][CH], s_wl[4][4][CH], s_w2[4][4][CH];
    __shared__ int s_off[4][65], s_j[4][64];
    __shared__ float s_z[4][16][16];
    __shared__ float4 s_wf[4][CH];  // W' in FP32 for a node redone by the FP32 pass
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = 1 + blockIdx.x * 4 + wv;
    if (i >= n_nodes) return;  // per wavefront: only wavefront-level LDS traffic below
    const int r = h.rank[i];
    if (r <= 0) {
        if (lane == 0) h.nflag[i] = 0;
        return;
    }
    const int G = r <= 4 ? 1 : r <= 8 ? 2 : r <= 16 ? 4 : r <= 32 ? 8 : 16;
    const int P = 16 / G;
    const int b = lane >> 2, u = lane & 3, g = b % G, ph = b / G;
    const float4 ci = h.cent[i];
    const int tr = 4 * g + u;
    const float4 xk = (tr < r) ? h.skelf[(size_t)i * RMAX + tr] : ci;
    v4f acc = {0.f, 0.f, 0.f, 0.f}, acc2 = {0.f, 0.f, 0.f, 0.f};
    const int a0 = h.adm_beg[i], a1 = h.adm_end[i], x0 = h.x_beg[i], x1 = h.x_end[i];
    const int na = a1 - a0, nm = na + (x1 - x0);
    for (int m0 = 0; m0 < nm; m0 += 64) {
        const int T = member_round<D>(h, m0, na, nm, a0, x0, lane, s_off[wv], s_j[wv]);
        const int nmem = min(64, nm - m0);
        for (int c0 = 0; c0 < T; c0 += CH) {
            // stage sources c0 .. c0 + CH - 1 (zero rows past the end)
            float4 xs, ws;
            stage_source<D>(h, ci, m0, na, nmem, c0 + lane, T, s_off[wv], s_j[wv], xs, ws);
            s_x[wv][lane] = xs;
            {
                const v4f wv4 = {ws.x, ws.y, ws.z, ws.w};
                const v4s hb = bf16x4(wv4);
                const v4s lb = bf16x4(wv4 - f32x4(hb));
                const v4s qb = bf16x4(wv4 * wv4);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    s_wh[wv][c][lane] = hb[c];
                    s_wl[wv][c][lane] = lb[c];
                    s_w2[wv][c][lane] = qb[c];
                }
            }
            __builtin_amdgcn_wave_barrier();
            const int nsteps = (min(CH, T - c0) + 4 * P - 1) / (4 * P);
            for (int t = 0; t < nsteps; t++) {
                const int sl = t * 4 * P + 4 * ph;  // sources sl .. sl+3 (< CH: 4P divides CH)
                v4f Kf;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float4 q = s_x[wv][sl + k];
                    Kf[k] = kern32(xk.x - q.x, xk.y - q.y, D == 3 ? xk.z - q.z : 0.f);
                }
                const v4s A = bf16x4(Kf), A2 = bf16x4(Kf * Kf);
                const v4s Bh = *reinterpret_cast<const v4s *>(&s_wh[wv][u][sl]);
                const v4s Bl = *reinterpret_cast<const v4s *>(&s_wl[wv][u][sl]);
                const v4s B2 = *reinterpret_cast<const v4s *>(&s_w2[wv][u][sl]);
                acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(A, Bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(A, Bl, acc, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(A2, B2, acc2, 0, 0, 0);
            }
            __builtin_amdgcn_wave_barrier();  // the next round overwrites the staging
        }
        __builtin_amdgcn_wave_barrier();
    }
    // sum the phases: blocks with the same row group are 4G lanes apart
    for (int o = 4 * G; o < 64; o <<= 1)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            acc[q] += __shfl_xor(acc[q], o, 64);
            acc2[q] += __shfl_xor(acc2[q], o, 64);
        }
    // lane (block b, column u) now holds z[4g + q][u], q = 0..3.  The scale an entry's
    // error is held to: |z| for the s column (u = 0); for a force column, max(|z|, L |z_0|)
    // with L the node's half-side, since the leaf expansion adds it next to
    // (y_t - c_i) z_0 with |y_t - c_i| up to L (a force column that cancels to near zero
    // is small against that term, not a reason to redo the node in FP32)
    const float L = ci.w;
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float z0 = __shfl(acc[q], lane & ~3, 64);
        const float scale = (u == 0) ? fabsf(acc[q]) : fmaxf(fabsf(acc[q]), L * fabsf(z0));
        if (4 * g + q < r && 0.001953125f * sqrtf(fmaxf(acc2[q], 0.f)) > tau * scale)
            bad = true;
    }
    const bool redo = __ballot(bad) != 0;
    if (lane == 0) h.nflag[i] = redo ? 1 : 0;  // every node's flag: no clearing pass
    if (redo) {
        __builtin_amdgcn_wave_barrier();
        couple_node_f32<D>(h, i, r, lane, s_x[wv], s_wf[wv], s_off[wv], s_j[wv]);
        return;
    }
    if (ph == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) s_z[wv][b][q * 4 + u] = acc[q];
    __builtin_amdgcn_wave_barrier();
    if (lane < r) {
        const float *zr = &s_z[wv][lane >> 2][(lane & 3) * 4];
        h.zc[(size_t)i * RMAX + lane] = make_float4(zr[0], zr[1], zr[2], zr[3]);
    }
}

// Downward sweep for the nodes of one level (>= 2): child c takes its rows of the
// parent's transfer matrix U_p.
template <int D>
__global__ void __launch_bounds__(256)
down_kernel(H2Dev h, int lb, int le)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    const int lane = threadIdx.x & 63;
    const int c = lb + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= le) return;
    const int p = h.parent[c];
    const int rp = h.rank[p], rc = h.rank[c];
    int row0 = 0;
    for (int s = h.child0[p]; s < c; s++) row0 += h.rank[s];
    const int RS = rank_stride(rc), G = 64 / RS;
    const int k = lane % RS, g = lane / RS;
    const float *Up = h.upool + h.uoff[p];
    const float4 *zp = h.zc + (size_t)p * RMAX;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < rc) {
        const float *urow = Up + (size_t)(row0 + k) * rp;
        for (int q = g; q < rp; q += G) {
            const float u = urow[q];
            const float4 z = zp[q];
            acc.x = fmaf(u, z.x, acc.x);
            acc.y = fmaf(u, z.y, acc.y);
            acc.z = fmaf(u, z.z, acc.z);
            acc.w = fmaf(u, z.w, acc.w);
        }
    }
    acc = group_sum(acc, RS);
    if (g == 0 && k < rc) {
        const float4 cc = h.cent[c], cp = h.cent[p];
        float4 z = h.zc[(size_t)c * RMAX + k];
        z.x += acc.x;
        z.y += acc.y + (cc.x - cp.x) * acc.x;
        z.z += acc.z + (cc.y - cp.y) * acc.x;
        z.w += acc.w + (cc.z - cp.z) * acc.x;
        h.zc[(size_t)c * RMAX + k] = z;
    }
}

// Mixed-precision downward sweep: the same update as down_kernel as a batched small GEMM
// on MFMA, one wavefront per child c: (r_c x r_p) rows of U_p times z_p (r_p x 4).  As in
// couple_mfma_kernel, v_mfma_f32_4x4x4bf16_1k runs 16 blocks: block b = lane / 4 takes
// rows 4g..4g+3 of the child (g = b mod G) against 4 parent ranks of phase b / G; lane u
// of a block supplies row 4g+u of U_p and column u of z_p.  Both operands are split into
// two BF16 terms (hi + lo, ~16 bits each); U_h z_h + U_h z_l + U_l z_h accumulate in FP32
// (the dropped U_l z_l term is below 2^-16 of the product).
template <int D>
__global__ void __launch_bounds__(256)
down_mfma_kernel(H2Dev h, int lb, int le)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    static_assert(RMAX <= 64, "one staged round of parent ranks");
    __shared__ __align__(8) short s_zh[4][4][64], s_zl[4][4][64];
    __shared__ float s_z[4][16][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lb + blockIdx.x * 4 + wv;
    if (c >= le) return;  // per wavefront: only wavefront-level LDS traffic below
    const int p = h.parent[c];
    const int rp = h.rank[p], rc = h.rank[c];
    if (rc <= 0) return;
    int row0 = 0;
    for (int s = h.child0[p]; s < c; s++) row0 += h.rank[s];
    // z_p staged as BF16 hi / lo columns (zero past r_p)
    {
        const float4 z = (lane < rp) ? h.zc[(size_t)p * RMAX + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        const v4f zv = {z.x, z.y, z.z, z.w};
        const v4s zh = bf16x4(zv), zl = bf16x4(zv - f32x4(zh));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            s_zh[wv][q][lane] = zh[q];
            s_zl[wv][q][lane] = zl[q];
        }
    }
    __builtin_amdgcn_wave_barrier();
    const int G = rc <= 4 ? 1 : rc <= 8 ? 2 : rc <= 16 ? 4 : rc <= 32 ? 8 : 16;
    const int P = 16 / G;
    const int b = lane >> 2, u = lane & 3, g = b % G, ph = b / G;
    const int tr = 4 * g + u;  // this lane's row of the child
    const float *urow = h.upool + h.uoff[p] + (size_t)(row0 + min(tr, rc - 1)) * rp;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 4 * ph; k0 < rp; k0 += 4 * P) {
        v4f a;
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] = (tr < rc && k0 + k < rp) ? urow[k0 + k] : 0.f;
        const v4s Ah = bf16x4(a), Al = bf16x4(a - f32x4(Ah));
        const v4s Bh = *reinterpret_cast<const v4s *>(&s_zh[wv][u][k0]);
        const v4s Bl = *reinterpret_cast<const v4s *>(&s_zl[wv][u][k0]);
        acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ah, Bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Ah, Bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Al, Bh, acc, 0, 0, 0);
    }
    // blocks of one row group in other phases are 4G lanes apart; phases past r_p
    // contributed nothing (their loop ran no step), so the sum is over every phase
    for (int o = 4 * G; o < 64; o <<= 1)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += __shfl_xor(acc[q], o, 64);
    if (ph == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) s_z[wv][b][q * 4 + u] = acc[q];
    __builtin_amdgcn_wave_barrier();
    if (lane < rc) {
        const float *zr = &s_z[wv][lane >> 2][(lane & 3) * 4];
        const float4 cc = h.cent[c], cp = h.cent[p];
        float4 z = h.zc[(size_t)c * RMAX + lane];
        z.x += zr[0];
        z.y += zr[1] + (cc.x - cp.x) * zr[0];
        z.z += zr[2] + (cc.y - cp.y) * zr[0];
        z.w += zr[3] + (cc.z - cp.z) * zr[0];
        h.zc[(size_t)c * RMAX + lane] = z;
    }
}

// Leaves: one wavefront per leaf, each lane one or two target points.  Far part from
// the leaf's coefficients, then the NEAR leaves (sources staged in LDS) and the W
// skeletons.  The leaf is in its own near list, so every target met itself once with
// K = 1: s -= 1 removes it (F gets 0 from it).  Results are scattered to the caller's
// point order.
template <int D>
__global__ void __launch_bounds__(64)
leaf_kernel(H2Dev h, const int *__restrict__ leaves, float *__restrict__ s_out,
            float *__restrict__ F_out)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    __shared__ VT<D> s_src[kLeafCap];
    __shared__ float4 s_sk[RMAX], s_w[RMAX];
    const int lane = threadIdx.x;
    const int i = leaves[blockIdx.x];
    const int b = h.nb[i];
    const int mtot = min(h.ne[i] - b, H2Cfg<D>::LEAF_CAP);
    const VT<D> *ys = (const VT<D> *)h.Ys;
    const float4 ci = h.cent[i];
    // targets in chunks of kLeafCap points (one chunk but for the oversized leaves of
    // the deepest level); each chunk walks the near and W lists
    for (int tb = 0; tb < mtot; tb += kLeafCap) {
    const int m = min(mtot - tb, kLeafCap) + tb;  // targets tb .. m-1 of this chunk
    const bool two = m - tb > 64;
    const int t0 = tb + lane, t1 = tb + lane + 64;
    VT<D> y0 = ys[b + min(t0, m - 1)];
    VT<D> y1 = ys[b + min(t1, m - 1)];
    float s0 = 0.f, s1 = 0.f;
    float f0x = 0.f, f0y = 0.f, f0z = 0.f, f1x = 0.f, f1y = 0.f, f1z = 0.f;
    // far field through the leaf's basis
    if (h.level[i] >= 1) {
        const int r = h.rank[i];
        const float *U = h.upool + h.uoff[i];
        const float4 *z = h.zc + (size_t)i * RMAX;
        float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
        for (int k = 0; k < r; k++) {
            const float4 zk = z[k];
            const float u0 = (t0 < m) ? U[(size_t)t0 * r + k] : 0.f;
            a0.x = fmaf(u0, zk.x, a0.x); a0.y = fmaf(u0, zk.y, a0.y);
            a0.z = fmaf(u0, zk.z, a0.z); a0.w = fmaf(u0, zk.w, a0.w);
            if (two) {
                const float u1 = (t1 < m) ? U[(size_t)t1 * r + k] : 0.f;
                a1.x = fmaf(u1, zk.x, a1.x); a1.y = fmaf(u1, zk.y, a1.y);
                a1.z = fmaf(u1, zk.z, a1.z); a1.w = fmaf(u1, zk.w, a1.w);
            }
        }
        s0 = a0.x;
        f0x = (y0.x - ci.x) * a0.x + a0.y;
        f0y = (y0.y - ci.y) * a0.x + a0.z;
        if constexpr (D == 3) f0z = (y0.z - ci.z) * a0.x + a0.w;
        s1 = a1.x;
        f1x = (y1.x - ci.x) * a1.x + a1.y;
        f1y = (y1.y - ci.y) * a1.x + a1.z;
        if constexpr (D == 3) f1z = (y1.z - ci.z) * a1.x + a1.w;
    }
    // near field: points <- points
    for (int e = h.nl_beg[i]; e < h.nl_end[i]; e++) {
        const int j = (int)(uint32_t)h.nl[e];
        const int mjt = min(h.ne[j] - h.nb[j], H2Cfg<D>::LEAF_CAP);
        for (int sb = 0; sb < mjt; sb += kLeafCap) {
        const int bj = h.nb[j] + sb;
        const int mj = min(mjt - sb, kLeafCap);
        __syncthreads();
        for (int u = lane; u < mj; u += 64) s_src[u] = ys[bj + u];
        __syncthreads();
        if (!two) {
            for (int u = 0; u < mj; u++) {
                const VT<D> q = s_src[u];
                const float dx = y0.x - q.x, dy = y0.y - q.y;
                float dz = 0.f;
                if constexpr (D == 3) dz = y0.z - q.z;
                const float K = kern32(dx, dy, dz);
                s0 += K;
                f0x = fmaf(K, dx, f0x);
                f0y = fmaf(K, dy, f0y);
                if constexpr (D == 3) f0z = fmaf(K, dz, f0z);
            }
        } else {
            for (int u = 0; u < mj; u++) {
                const VT<D> q = s_src[u];
                const float dx0 = y0.x - q.x, dy0 = y0.y - q.y;
                const float dx1 = y1.x - q.x, dy1 = y1.y - q.y;
                float dz0 = 0.f, dz1 = 0.f;
                if constexpr (D == 3) { dz0 = y0.z - q.z; dz1 = y1.z - q.z; }
                const float K0 = kern32(dx0, dy0, dz0);
                const float K1 = kern32(dx1, dy1, dz1);
                s0 += K0;
                f0x = fmaf(K0, dx0, f0x);
                f0y = fmaf(K0, dy0, f0y);
                s1 += K1;
                f1x = fmaf(K1, dx1, f1x);
                f1y = fmaf(K1, dy1, f1y);
                if constexpr (D == 3) { f0z = fmaf(K0, dz0, f0z); f1z = fmaf(K1, dz1, f1z); }
            }
        }
        }  // source chunks
    }
    // W: points <- skeleton of a finer node
    for (int e = h.w_beg[i]; e < h.w_end[i]; e++) {
        const int d = (int)(uint32_t)h.wl[e];
        const int rd = h.rank[d];
        __syncthreads();
        if (lane < rd) {
            s_sk[lane] = h.skelf[(size_t)d * RMAX + lane];
            s_w[lane] = h.wc[(size_t)d * RMAX + lane];
        }
        __syncthreads();
        const float4 cdn = h.cent[d];
        for (int q = 0; q < rd; q++) {
            const float4 x = s_sk[q], w = s_w[q];
            {
                float dz = 0.f;
                if constexpr (D == 3) dz = y0.z - x.z;
                const float K = kern32(y0.x - x.x, y0.y - x.y, dz);
                const float a = K * w.x;
                s0 += a;
                f0x += (y0.x - cdn.x) * a - K * w.y;
                f0y += (y0.y - cdn.y) * a - K * w.z;
                if constexpr (D == 3) f0z += (y0.z - cdn.z) * a - K * w.w;
            }
            if (two) {
                float dz = 0.f;
                if constexpr (D == 3) dz = y1.z - x.z;
                const float K = kern32(y1.x - x.x, y1.y - x.y, dz);
                const float a = K * w.x;
                s1 += a;
                f1x += (y1.x - cdn.x) * a - K * w.y;
                f1y += (y1.y - cdn.y) * a - K * w.z;
                if constexpr (D == 3) f1z += (y1.z - cdn.z) * a - K * w.w;
            }
        }
    }
    if (t0 < m) {
        const int o = h.perm[b + t0];
        s_out[o] = s0 - 1.0f;
        F_out[(size_t)o * D] = f0x;
        F_out[(size_t)o * D + 1] = f0y;
        if constexpr (D == 3) F_out[(size_t)o * D + 2] = f0z;
    }
    if (two && t1 < m) {
        const int o = h.perm[b + t1];
        s_out[o] = s1 - 1.0f;
        F_out[(size_t)o * D] = f1x;
        F_out[(size_t)o * D + 1] = f1y;
        if constexpr (D == 3) F_out[(size_t)o * D + 2] = f1z;
    }
    }  // target chunks
}

int launch_couple(int D, const H2Dev &h, int n_nodes, int mixed, float tau, hipStream_t st)
{
    const int g = (n_nodes - 1 + 3) / 4;
    if (g <= 0) return 0;
    if (mixed) {
        if (D == 2)
            hipLaunchKernelGGL(couple_mfma_kernel<2>, dim3(g), dim3(256), 0, st, h, n_nodes, tau);
        else
            hipLaunchKernelGGL(couple_mfma_kernel<3>, dim3(g), dim3(256), 0, st, h, n_nodes, tau);
    } else if (D == 2) {
        hipLaunchKernelGGL(couple_kernel<2>, dim3(g), dim3(256), 0, st, h, n_nodes);
    } else {
        hipLaunchKernelGGL(couple_kernel<3>, dim3(g), dim3(256), 0, st, h, n_nodes);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_down(int D, const H2Dev &h, int lb, int le, int mixed, hipStream_t st)
{
    const int g = (le - lb + 3) / 4;
    if (g <= 0) return 0;
    if (mixed && D == 2)
        hipLaunchKernelGGL(down_mfma_kernel<2>, dim3(g), dim3(256), 0, st, h, lb, le);
    else if (mixed)
        hipLaunchKernelGGL(down_mfma_kernel<3>, dim3(g), dim3(256), 0, st, h, lb, le);
    else if (D == 2)
        hipLaunchKernelGGL(down_kernel<2>, dim3(g), dim3(256), 0, st, h, lb, le);
    else
        hipLaunchKernelGGL(down_kernel<3>, dim3(g), dim3(256), 0, st, h, lb, le);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_leaf_eval(int D, const H2Dev &h, int n_leaves, float *s, float *F, hipStream_t st)
{
    if (n_leaves <= 0) return 0;
    if (D == 2)
        hipLaunchKernelGGL(leaf_kernel<2>, dim3(n_leaves), dim3(64), 0, st, h, h.leaves, s, F);
    else
        hipLaunchKernelGGL(leaf_kernel<3>, dim3(n_leaves), dim3(64), 0, st, h, h.leaves, s, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsne
