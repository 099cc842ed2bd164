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

// One wavefront per target node (all non-root nodes); lanes are (group g, row k): row k
// of the node's skeleton, groups split the source list.
template <int D>
__global__ void __launch_bounds__(256)
couple_kernel(H2Dev h, int n_nodes)
{
    constexpr int RMAX = H2Cfg<D>::RMAX;
    const int lane = threadIdx.x & 63;
    const int i = 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n_nodes) return;
    const int r = h.rank[i];
    const int RS = rank_stride(r), G = 64 / RS;
    const int k = lane % RS, g = lane / RS;
    const float4 ci = h.cent[i];
    const float4 xk = (k < r) ? h.skelf[(size_t)i * RMAX + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int a0 = h.adm_beg[i], a1 = h.adm_end[i];
    for (int e = a0 + g; e < a1; e += G) {
        const int j = (int)(uint32_t)h.adm[e];
        const int rj = h.rank[j];
        const float4 cj = h.cent[j];
        const float dcx = ci.x - cj.x, dcy = ci.y - cj.y, dcz = ci.z - cj.z;
        const float4 *sk = h.skelf + (size_t)j * RMAX;
        const float4 *wj = h.wc + (size_t)j * RMAX;
        for (int q = 0; q < rj; q++) {
            const float4 xl = sk[q];
            const float4 w = wj[q];
            const float K = kern32(xk.x - xl.x, xk.y - xl.y, D == 3 ? xk.z - xl.z : 0.f);
            const float a = K * w.x;
            acc.x += a;
            acc.y = fmaf(dcx, a, fmaf(-K, w.y, acc.y));
            acc.z = fmaf(dcy, a, fmaf(-K, w.z, acc.z));
            if constexpr (D == 3) acc.w = fmaf(dcz, a, fmaf(-K, w.w, acc.w));
        }
    }
    const int x0 = h.x_beg[i], x1 = h.x_end[i];
    for (int e = x0 + g; e < x1; e += G) {
        const int j = (int)(uint32_t)h.xl[e];
        const VT<D> *ys = (const VT<D> *)h.Ys;
        for (int u = h.nb[j]; u < h.ne[j]; u++) {
            const VT<D> y = ys[u];
            float dz = 0.f;
            if constexpr (D == 3) dz = xk.z - y.z;
            const float K = kern32(xk.x - y.x, xk.y - y.y, dz);
            acc.x += K;
            acc.y = fmaf(K, ci.x - y.x, acc.y);
            acc.z = fmaf(K, ci.y - y.y, acc.z);
            if constexpr (D == 3) acc.w = fmaf(K, ci.z - y.z, acc.w);
        }
    }
    acc = group_sum(acc, RS);
    if (g == 0 && k < r) h.zc[(size_t)i * RMAX + k] = acc;
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
    const int m = min(h.ne[i] - b, kLeafCap);
    const VT<D> *ys = (const VT<D> *)h.Ys;
    const float4 ci = h.cent[i];
    const bool two = m > 64;
    const int t0 = lane, t1 = lane + 64;
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
        const int bj = h.nb[j];
        const int mj = min(h.ne[j] - bj, kLeafCap);
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
}

int launch_couple(int D, const H2Dev &h, int n_nodes, int mixed, hipStream_t st)
{
    (void)mixed;
    const int g = (n_nodes - 1 + 3) / 4;
    if (g <= 0) return 0;
    if (D == 2)
        hipLaunchKernelGGL(couple_kernel<2>, dim3(g), dim3(256), 0, st, h, n_nodes);
    else
        hipLaunchKernelGGL(couple_kernel<3>, dim3(g), dim3(256), 0, st, h, n_nodes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_down(int D, const H2Dev &h, int lb, int le, int mixed, hipStream_t st)
{
    (void)mixed;
    const int g = (le - lb + 3) / 4;
    if (g <= 0) return 0;
    if (D == 2)
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
