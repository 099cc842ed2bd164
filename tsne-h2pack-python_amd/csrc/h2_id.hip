// h2_id.hip -- H2 rebuild, part 2: proxy-point interpolative decomposition (ID) of every
// node, fused with the upward pass of the product.
//
// Node i (level l >= 1) has m candidate points: its own points if it is a leaf, else
// the skeleton points of its children.  With the proxy set P of its level translated to
// its centre, A = K(candidates, P) (m x p) is column-normalised and row-ID'd by
// column-pivoted modified Gram-Schmidt on A^T in FP64:  A ~= U A[J, :], |J| = r, where
// the pivoting stops once the largest residual column norm is below tol (relative to
// the first pivot).  The Gram-Schmidt dot products are the entries of R: each candidate
// stores its column of R (FP64) in the node's region of the U pool as the pivoting runs,
// and U = R11^{-1} R is formed from them, so R is never recomputed and only the current
// row of Q is kept.  U (m x r, FP32, written over the R entries) and the skeleton J are
// stored; then the upward
// coefficients w_i = U^T X_i are formed, X_i holding one row [1, y - c_i] per leaf point
// or, for an internal node, the children's coefficients moved to c_i
// ([w0, wY + (c_child - c_i) w0]).
//
// A node with a single child keeps the child's skeleton (U = I): the child's far region
// (outside 3 child half-sides of its centre) contains the parent's, so the child basis
// already represents every interaction the parent's does.
//
// Reference counterpart: the U / J construction of H2P_build in H2Pack (QR_REL_NRM ID
// against proxy points; src/tsne_h2pack.c:818, :908-919) and the upward sweep of the
// fused product (src/tsne_h2_matmul.c:309-334).  One workgroup per node, one thread per
// candidate column.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "h2_internal.h"
#include "proxy.h"

namespace tsne {

namespace {

template <int D> using VT = typename Vec<D>::T;

__device__ __forceinline__ double kern64(const double *a, const double *b, int D)
{
    double r2 = 1.0;
    for (int d = 0; d < D; d++) {
        const double df = a[d] - b[d];
        r2 += df * df;
    }
    // hardware reciprocal refined by one Newton step (full FP64 accuracy, no IEEE divide)
    double w = __builtin_amdgcn_rcp(r2);
    w = fma(w, fma(-r2, w, 1.0), w);
    return w * w;
}

// Reduce-scatter of column sums of squares over the wavefront: lane j returns the sum over
// the 64 lanes of a[C0 + j]^2 (columns at or past NA count zero).  At each of six steps a
// lane keeps half of its partial sums and adds the partner's copy of that half: 63
// independent shuffles for 64 columns instead of a 6-step chain per column.
template <int C0, int NA, typename T>
__device__ __forceinline__ double colsq_scatter64(const T (&a)[NA], int lane)
{
    double w1[32], w2[16], w3[8], w4[4], w5[2];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4, b1 = lane & 2,
               b0 = lane & 1;
#pragma unroll
    for (int j = 0; j < 32; j++) {
        const double lo = (C0 + j < NA) ? (double)a[C0 + j] * a[C0 + j] : 0.0;
        const double hi = (C0 + j + 32 < NA) ? (double)a[C0 + j + 32] * a[C0 + j + 32] : 0.0;
        w1[j] = (b5 ? hi : lo) + __shfl_xor(b5 ? lo : hi, 32, 64);
    }
#pragma unroll
    for (int j = 0; j < 16; j++)
        w2[j] = (b4 ? w1[j + 16] : w1[j]) + __shfl_xor(b4 ? w1[j] : w1[j + 16], 16, 64);
#pragma unroll
    for (int j = 0; j < 8; j++)
        w3[j] = (b3 ? w2[j + 8] : w2[j]) + __shfl_xor(b3 ? w2[j] : w2[j + 8], 8, 64);
#pragma unroll
    for (int j = 0; j < 4; j++)
        w4[j] = (b2 ? w3[j + 4] : w3[j]) + __shfl_xor(b2 ? w3[j] : w3[j + 4], 4, 64);
#pragma unroll
    for (int j = 0; j < 2; j++)
        w5[j] = (b1 ? w4[j + 2] : w4[j]) + __shfl_xor(b1 ? w4[j] : w4[j + 2], 2, 64);
    return (b0 ? w5[1] : w5[0]) + __shfl_xor(b0 ? w5[0] : w5[1], 1, 64);
}

// Pivot key of a candidate: the float bits of its residual norm (monotonic for norms >= 0)
// with the low 9 bits replaced by 511 - t, so that the unsigned maximum picks the largest
// norm and, among norms equal to within 2^-14, the smallest t (t < 512).  0: out of play.
__device__ __forceinline__ unsigned pivot_key(double nrm, int t, bool live)
{
    if (!live) return 0u;
    const unsigned b = __float_as_uint((float)nrm);
    return (b & ~0x1FFu) | (unsigned)(511 - t);
}

// Unsigned maximum over the wavefront, in every lane: DPP within each row of 16 lanes
// (quad swaps, then rotations by 4 and 8), then the four row maxima through scalar
// registers.  Every lane of the wavefront must be active.
__device__ __forceinline__ unsigned wave_max_u32(unsigned k)
{
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0xB1, 0xF, 0xF, false));  // lanes ^1
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x4E, 0xF, 0xF, false));  // lanes ^2
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x124, 0xF, 0xF, false)); // row_ror:4
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, 0x128, 0xF, 0xF, false)); // row_ror:8
    const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)k, 0);
    const unsigned r1 = (unsigned)__builtin_amdgcn_readlane((int)k, 16);
    const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)k, 32);
    const unsigned r3 = (unsigned)__builtin_amdgcn_readlane((int)k, 48);
    return max(max(r0, r1), max(r2, r3));
}

}  // namespace

// ID + upward pass of one node, by one workgroup of NT threads; nodes with at most
// m_lo candidates, or (below the maximum size class) more than NT, are left to another
// size class.
// PQ proxy slots are processed (PQ <= the table's PMAX; at least the largest proxy count
// of any level).  T: arithmetic of the pivoted Gram-Schmidt (candidate columns, the Q row,
// dot products and residual norms).  double everywhere by default; float for the mixed
// path at ID tolerances >= 1e-3, where FP32 rounding (~6e-8) is far below the residual
// ratios the pivoting compares (>= tol), and half the LDS bytes per broadcast Q read.  The
// kernel values, column scales, R entries, triangular solve and U's coefficients stay FP64.
template <int D, int NT, int PQ, typename T>
__device__ __forceinline__ void id_up_node(const H2Dev &h, const int node, int m_lo, double tol)
{
    constexpr int PMAX = PQ, RMAX = H2Cfg<D>::RMAX;
    constexpr int PSTRIDE = H2Cfg<D>::PMAX;  // row stride of the proxy table
    static_assert(PMAX % 4 == 0 && PMAX <= PSTRIDE, "proxy slots");
    constexpr int NW = NT / 64;
    __shared__ float4 s_xin[NT];
    __shared__ double s_prox[PMAX][D];
    __shared__ double s_part[NW][PMAX];
    __shared__ double s_scale[PMAX];
    __shared__ T s_Q[2][PMAX];  // Q row of the current step (two, alternating)
    __shared__ float s_U[NT][RMAX];
    __shared__ double s_R[RMAX][RMAX];
    __shared__ int s_piv[RMAX];
    __shared__ unsigned s_key[2][NW];  // wavefront pivot keys (two steps, alternating)
    __shared__ T s_bv;
    __shared__ int s_base;

    constexpr int NCH = H2Cfg<D>::NCH;
    const int t = threadIdx.x;
    const int lane = t & 63, wid = t >> 6;
    // node record and, independent of it, the proxy set of the node's level: all these
    // loads are issued together
    const int ch0 = h.child0[node];
    const bool leafn = ch0 < 0;
    const int l = h.level[node];
    const float4 c = h.cent[node];
    const RootInfo root = *h.root;
    const int kq = min(max(root.k0 - l, 0), kProxyLevels - 1);
    const int p = min(h.proxy_cnt[kq], PMAX);
    // proxies t, t + NT, ... (relative to the centre) staged by this thread
    constexpr int PPT = (PMAX + NT - 1) / NT;
    double px[PPT][3] = {};
#pragma unroll
    for (int u = 0; u < PPT; u++)
        if (t + u * NT < p)
            for (int d = 0; d < D; d++)
                px[u][d] = h.proxy_pts[((size_t)kq * PSTRIDE + t + u * NT) * D + d];
    int nb = 0, mt = 0, nc = 0;
    int rk[NCH];
    if (leafn) {
        nb = h.nb[node];
        mt = h.ne[node] - nb;
    } else {
        nc = h.nchild[node];
#pragma unroll
        for (int q = 0; q < NCH; q++) rk[q] = (q < nc) ? h.rank[ch0 + q] : 0;
#pragma unroll
        for (int q = 0; q < NCH; q++) mt += rk[q];
    }
    // size class: this instantiation takes the nodes with m_lo < m <= NT candidates
    if (mt <= m_lo || (mt > NT && NT < H2Cfg<D>::MCAP)) return;
    const int m = min(mt, NT);
    // U of the node: a leaf's rows are its points, so its block sits at its first sorted
    // point (m x RMAX floats); an internal node reserves m x RMAX floats after the n x RMAX
    // of the leaves (the reservation is issued now, its latency hidden by the ID)
    // (region blockIdx mod kPoolParts of the internal pool, so that the reservations of
    // one launch spread over kPoolParts counters)
    // Regions are m x 2 RMAX floats: the FP64 R entries (r x m) while the ID runs, then
    // U (m x r floats) written over them.  Offsets are even (8-byte aligned doubles).
    const int part = blockIdx.x % kPoolParts;
    const long long part_cap = pool_part_cap(h, RMAX);
    if (t == 0)
        s_base = leafn ? nb * 2 * RMAX
                       : (int)(2LL * h.n * RMAX + part * part_cap +
                               atomicAdd(&h.cnt[CNT_POOL + part * kCntSlot], 2 * m * RMAX));
    const double cd[3] = {(double)c.x, (double)c.y, (double)c.z};

    // ---- candidates and the rows of X
    double my[3] = {0.0, 0.0, 0.0};
    float4 xin = make_float4(0.f, 0.f, 0.f, 0.f);
    if (leafn) {
        if (t < m) {
            const VT<D> y = ((const VT<D> *)h.Ys)[nb + t];
            my[0] = y.x;
            my[1] = y.y;
            if constexpr (D == 3) my[2] = y.z;
            xin = make_float4(1.f, y.x - c.x, y.y - c.y, 0.f);
            if constexpr (D == 3) xin.w = y.z - c.z;
        }
    } else {
        // this thread's child and skeleton slot, from the children's ranks
        int cj = -1, k = 0, acc = 0;
#pragma unroll
        for (int q = 0; q < NCH; q++) {
            if (t >= acc && t < acc + rk[q]) {
                cj = ch0 + q;
                k = t - acc;
            }
            acc += rk[q];
        }
        if (cj >= 0) {
            for (int d = 0; d < D; d++) my[d] = h.skel[((size_t)cj * RMAX + k) * D + d];
            const float4 w = h.wc[(size_t)cj * RMAX + k];
            const float4 cc = h.cent[cj];
            xin = make_float4(w.x, w.y + (cc.x - c.x) * w.x, w.z + (cc.y - c.y) * w.x,
                              w.w + (cc.z - c.z) * w.x);
        }
        if (nc == 1) {
            // single child: same skeleton, U = I, coefficients moved to this centre
            __syncthreads();
            const int base = s_base;
            const bool fits = (long long)base + 2LL * m * RMAX <=
                              2LL * h.n * RMAX + (part + 1) * part_cap;
            if (t == 0) {
                if (!fits) atomicOr(&h.flags[FLAG_POOL_OVF], 1);
                h.rank[node] = m;
                h.uoff[node] = base;
            }
            if (t < m) {
                if (fits)
                    for (int q = 0; q < m; q++) h.upool[(size_t)base + (size_t)t * m + q] = (q == t) ? 1.f : 0.f;
                for (int d = 0; d < D; d++) h.skel[((size_t)node * RMAX + t) * D + d] = my[d];
                h.skelf[(size_t)node * RMAX + t] =
                    make_float4((float)my[0], (float)my[1], (float)my[2], 0.f);
                h.wc[(size_t)node * RMAX + t] = xin;
            }
            return;
        }
    }
    s_xin[t] = xin;

    // ---- proxies of this level, translated to the centre.  Everything below runs over
    // all PMAX proxy slots and all RMAX rank slots without branches: unused slots hold
    // zeros (scale 0, zero Q rows, identity R rows), which keeps the FP64 pipelines full.
#pragma unroll
    for (int u = 0; u < PPT; u++)
        if (t + u * NT < PMAX)
            for (int d = 0; d < D; d++) s_prox[t + u * NT][d] = cd[d] + px[u][d];
    for (int e = t; e < RMAX * RMAX; e += NT) s_R[e / RMAX][e % RMAX] = (e / RMAX == e % RMAX) ? 1.0 : 0.0;
    __syncthreads();

    // ---- this thread's column of A^T, then column norms over the candidates
    T a[PMAX];
#pragma unroll
    for (int q = 0; q < PMAX; q++) a[q] = (t < m) ? (T)kern64(my, s_prox[q], D) : T(0);
    if constexpr (D == 2) {
        // reduce-scatter over the wavefront: at each step a lane keeps half of its
        // partial column sums and adds the partner's copy of that half, so after six
        // steps lane j holds the wavefront's sum for column j (mod 32 when PMAX <= 32):
        // 63 independent shuffles instead of a 6-step chain per column
        static_assert(PMAX <= 64, "2D column count");
        double w1[32], w2[16], w3[8], w4[4], w5[2];
        const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4, b1 = lane & 2,
                   b0 = lane & 1;
#pragma unroll
        for (int j = 0; j < 32; j++) {
            if constexpr (PMAX <= 32) {
                const double v = (j < PMAX) ? (double)a[j] * a[j] : 0.0;
                w1[j] = v + __shfl_xor(v, 32, 64);
            } else {
                const double lo = (double)a[j] * a[j];
                const double hi = (j + 32 < PMAX) ? (double)a[j + 32] * a[j + 32] : 0.0;
                w1[j] = (b5 ? hi : lo) + __shfl_xor(b5 ? lo : hi, 32, 64);
            }
        }
#pragma unroll
        for (int j = 0; j < 16; j++)
            w2[j] = (b4 ? w1[j + 16] : w1[j]) + __shfl_xor(b4 ? w1[j] : w1[j + 16], 16, 64);
#pragma unroll
        for (int j = 0; j < 8; j++)
            w3[j] = (b3 ? w2[j + 8] : w2[j]) + __shfl_xor(b3 ? w2[j] : w2[j + 8], 8, 64);
#pragma unroll
        for (int j = 0; j < 4; j++)
            w4[j] = (b2 ? w3[j + 4] : w3[j]) + __shfl_xor(b2 ? w3[j] : w3[j + 4], 4, 64);
#pragma unroll
        for (int j = 0; j < 2; j++)
            w5[j] = (b1 ? w4[j + 2] : w4[j]) + __shfl_xor(b1 ? w4[j] : w4[j + 2], 2, 64);
        const double w6 = (b0 ? w5[1] : w5[0]) + __shfl_xor(b0 ? w5[0] : w5[1], 1, 64);
        if (lane < PMAX) s_part[wid][lane] = w6;
        (void)b5;
    } else {
        // 3D: columns 0-63, then 64-95, each reduce-scattered over the wavefront
        static_assert(PMAX <= 128, "3D column count");
        const double c0 = colsq_scatter64<0, PMAX>(a, lane);
        if (lane < PMAX) s_part[wid][lane] = c0;
        if constexpr (PMAX > 64) {
            const double c1 = colsq_scatter64<64, PMAX>(a, lane);
            if (lane + 64 < PMAX) s_part[wid][64 + lane] = c1;
        }
    }
    __syncthreads();
    for (int q = t; q < PMAX; q += NT) {
        double s2 = 0.0;
        for (int w = 0; w < NW; w++) s2 += s_part[w][q];
        s_scale[q] = (q < p && s2 > 0.0) ? 1.0 / sqrt(s2) : 0.0;
    }
    __syncthreads();
    T nrm;
    {
        T n0 = 0, n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
        for (int q = 0; q < PMAX; q += 4) {
            a[q] *= (T)s_scale[q];
            a[q + 1] *= (T)s_scale[q + 1];
            a[q + 2] *= (T)s_scale[q + 2];
            a[q + 3] *= (T)s_scale[q + 3];
            n0 = fma(a[q], a[q], n0);
            n1 = fma(a[q + 1], a[q + 1], n1);
            n2 = fma(a[q + 2], a[q + 2], n2);
            n3 = fma(a[q + 3], a[q + 3], n3);
        }
        nrm = (t < m) ? (n0 + n1) + (n2 + n3) : T(-1);
    }

    // ---- column-pivoted MGS.  R[k][t] stays in thread t's registers (2D, <= 32 proxy
    // slots) or goes to rpool[k * m + t] (the node's region; an internal node whose reservation
    // overflowed the pool writes nothing, and the build is redone with a larger pool)
    const int rbase = s_base;
    const bool rfits = leafn || (long long)rbase + 2LL * m * RMAX <= 2LL * h.n * RMAX + (part + 1) * part_cap;
    double *rpool = reinterpret_cast<double *>(h.upool + rbase);
    bool selected = false;
    T nrm0 = 0;
    int r = 0;
    const int rlim = min(RMAX, min(m, p));
    const T tol2 = (T)(tol * tol);
    // 2D with up to 32 proxy slots: each candidate keeps its column of R in registers (the
    // step loop is unrolled, so every index is static); otherwise (3D, 48 proxy slots),
    // where the columns already fill the registers, it stores them in the pool region and
    // reads them back for the solve
    constexpr bool RREG = (D == 2 && PMAX <= 32);
    double rreg[RREG ? RMAX : 1];
#pragma unroll
    for (int k = 0; k < (RREG ? RMAX : 1); k++) rreg[k] = 0.0;
    // one pivoting step; false once the pivoting stops
    auto gs_step = [&](int k, double &rk) __attribute__((always_inline)) -> bool {
        // pivot: the candidate of largest residual norm (ties within 2^-14 to the smaller
        // index), by one 32-bit key maximum over the workgroup
        unsigned key = wave_max_u32(pivot_key((double)nrm, t, !selected && t < m));
        if (NW > 1) {
            if (lane == 0) s_key[k & 1][wid] = key;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < NW; w++) key = max(key, s_key[k & 1][w]);
        }
        const int bi = key ? 511 - (int)(key & 0x1FFu) : -1;
        // the pivot publishes its exact norm and, speculatively, its Q row (unused if the
        // pivoting stops here)
        if (t == bi) {
            s_bv = nrm;
            const T inv = (T)(1.0 / sqrt((double)nrm));
#pragma unroll
            for (int q = 0; q < PMAX; q++) s_Q[k & 1][q] = a[q] * inv;
            s_piv[k] = t;
        }
        __syncthreads();
        const T bv = key ? s_bv : T(0);
        if (k == 0) nrm0 = bv;
        if (k >= rlim || !(bv > tol2 * nrm0) || !(bv > T(0))) return false;
        if (t == bi) {
            selected = true;
            s_R[k][k] = sqrt((double)bv);  // R[k][k]
        }
        if (!selected && t < m) {
            // the Q row is read from LDS once into registers where it fits beside the
            // column (2D, and 3D in FP32), else once per pass
            constexpr bool QREG = PMAX * sizeof(T) <= 384;
            T qv[QREG ? PMAX : 1];
            if constexpr (QREG) {
#pragma unroll
                for (int q = 0; q < PMAX; q++) qv[q] = s_Q[k & 1][q];
            }
            auto Qk = [&](int q) -> T {
                if constexpr (QREG) return qv[q];
                else return s_Q[k & 1][q];
            };
            T d0 = 0, d1 = 0, d2 = 0, d3 = 0;
#pragma unroll
            for (int q = 0; q < PMAX; q += 4) if (D == 2 || q < p) {  // 3D: skip empty slots (register pressure)
                d0 = fma(Qk(q), a[q], d0);
                d1 = fma(Qk(q + 1), a[q + 1], d1);
                d2 = fma(Qk(q + 2), a[q + 2], d2);
                d3 = fma(Qk(q + 3), a[q + 3], d3);
            }
            const T dot = (d0 + d1) + (d2 + d3);  // R[k][t]
            if constexpr (RREG) rk = (double)dot;
            else if (rfits) rpool[(size_t)k * m + t] = (double)dot;
            T n0 = 0, n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
            for (int q = 0; q < PMAX; q += 4) if (D == 2 || q < p) {  // 3D: skip empty slots (register pressure)
                a[q] = fma(-dot, Qk(q), a[q]);
                a[q + 1] = fma(-dot, Qk(q + 1), a[q + 1]);
                a[q + 2] = fma(-dot, Qk(q + 2), a[q + 2]);
                a[q + 3] = fma(-dot, Qk(q + 3), a[q + 3]);
                n0 = fma(a[q], a[q], n0);
                n1 = fma(a[q + 1], a[q + 1], n1);
                n2 = fma(a[q + 2], a[q + 2], n2);
                n3 = fma(a[q + 3], a[q + 3], n3);
            }
            nrm = (n0 + n1) + (n2 + n3);
        }
        r = k + 1;
        if (k == RMAX - 1 && t == 0) atomicOr(&h.flags[FLAG_RANK_CAP], 1);
        return true;
    };
    if constexpr (RREG) {
#pragma unroll
        for (int k = 0; k < RMAX; k++)
            if (!gs_step(k, rreg[k])) break;
    } else {
        double unused = 0.0;
        for (int k = 0; k < RMAX; k++)
            if (!gs_step(k, unused)) break;
    }
    // The rest runs over RB rank slots, RB the smallest bucket (RMAX/4, RMAX/2, RMAX)
    // that holds r: a uniform branch per workgroup, each bucket fully unrolled.
    int kk = -1;
    if (selected)
        for (int k = 0; k < r; k++)
            if (s_piv[k] == t) kk = k;
    auto tail = [&](auto rb) {
        constexpr int RB = decltype(rb)::value;
        // this thread's column of R: rows 0..r-1 of a candidate, rows 0..kk of a pivot
        double rt[RB];
        const int rows = (kk >= 0) ? kk : r;
#pragma unroll
        for (int k = 0; k < RB; k++)
            if constexpr (RREG) rt[k] = (t < m && k < rows) ? rreg[k] : 0.0;
            else rt[k] = (t < m && k < rows && rfits) ? rpool[(size_t)k * m + t] : 0.0;
        if (kk >= 0) {
#pragma unroll
            for (int j = 0; j < RB; j++)
                if (j < kk) s_R[j][kk] = rt[j];  // the diagonal was set at selection
        }
        __syncthreads();  // s_R complete; every R entry of the pool read
        if (t < RB) s_part[0][t] = 1.0 / s_R[t][t];  // s_part reused: inverse diagonal
        __syncthreads();
        // ---- interpolation coefficients x_t = R11^{-1} R[:, t]
        if (kk >= 0) {
#pragma unroll
            for (int k = 0; k < RB; k++) rt[k] = (k == kk) ? 1.0 : 0.0;
        } else if (t < m) {
#pragma unroll
            for (int k = RB - 1; k >= 0; k--) {
                double v = rt[k];
#pragma unroll
                for (int j = k + 1; j < RB; j++) v = fma(-s_R[k][j], rt[j], v);
                rt[k] = v * s_part[0][k];
            }
        }
#pragma unroll
        for (int k = 0; k < RB; k++) s_U[t][k] = (float)rt[k];
    };
    if (r <= RMAX / 4)
        tail(std::integral_constant<int, RMAX / 4>{});
    else if (r <= RMAX / 2)
        tail(std::integral_constant<int, RMAX / 2>{});
    else
        tail(std::integral_constant<int, RMAX>{});

    // ---- store U (pool), skeleton, and the upward coefficients
    if (t == 0) {
        const int base = s_base;
        if (!rfits) atomicOr(&h.flags[FLAG_POOL_OVF], 1);
        h.rank[node] = r;
        h.uoff[node] = base;
    }
    if (kk >= 0) {
        for (int d = 0; d < D; d++) h.skel[((size_t)node * RMAX + kk) * D + d] = my[d];
        h.skelf[(size_t)node * RMAX + kk] =
            make_float4((float)my[0], (float)my[1], (float)my[2], 0.f);
    }
    __syncthreads();
    const int base = s_base;
    if (rfits)
        for (int e = t; e < m * r; e += NT) h.upool[(size_t)base + e] = s_U[e / r][e % r];
    if (t < r) {
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int u = 0; u < m; u++) {
            const float uk = s_U[u][t];
            const float4 x = s_xin[u];
            w.x = fmaf(uk, x.x, w.x);
            w.y = fmaf(uk, x.y, w.y);
            w.z = fmaf(uk, x.z, w.z);
            w.w = fmaf(uk, x.w, w.w);
        }
        h.wc[(size_t)node * RMAX + t] = w;
    }
}

// FP64 residual norms of the tiled 3D ID: downdated per step (true) or recomputed from the
// updated column (TSNE_ID_NO_DOWNDATE at compile time).  2D recomputes them (the step is
// short there, and the downdated form costs the 64-candidate class registers).
#ifdef TSNE_ID_NO_DOWNDATE
constexpr bool kDowndate = false;
#else
constexpr bool kDowndate = true;
#endif

// A DPP move of a 64-bit value (two 32-bit halves); CTRL a DPP control within a row.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the QG (4 or 8) lanes QG i .. QG i + QG - 1: lanes ^1 and ^2 (quad
// permutations), then for eight lanes the other quad through the half-row mirror (lane
// j <-> 7 - j).  Every lane of the group gets the same bits (each step adds the same two
// operands in either order).  All lanes of the group must be active.
template <int QG, typename T>
__device__ __forceinline__ T sum_group(T v)
{
    static_assert(QG == 4 || QG == 8, "group width");
    if constexpr (std::is_same<T, double>::value) {
        v += dpp_f64<0xB1>(v);
        v += dpp_f64<0x4E>(v);
        if constexpr (QG == 8) v += dpp_f64<0x141>(v);
    } else {
        v += dpp_f32<0xB1>(v);
        v += dpp_f32<0x4E>(v);
        if constexpr (QG == 8) v += dpp_f32<0x141>(v);
    }
    return v;
}

// ID + upward pass of one node by one workgroup of NT threads, the candidate columns
// split into tiles: thread t holds the CT x QT block of A for candidates g CT .. g CT +
// CT - 1 (g = t / QG) and proxy slots qg QT .. qg QT + QT - 1 (qg = t % QG, QT = PQ / QG).
// A pivot step's dot products and residual norms are then sums over the QG lanes of a
// group (DPP), each thread reads QT entries of the broadcast Q row instead of PQ and
// holds CT QT values of A instead of PQ (in 3D, QG = 8 and PQ = 96: no register spills).
// The candidates' set-up, the triangular solve and the stores run one candidate per
// thread.  Same arithmetic as id_up_node (FP64 kernel values, column scales, R entries and
// solve; T for the Gram-Schmidt); sums are taken in another order.  PQ proxy slots are
// processed (at least the largest proxy count of any level).  Nodes with m_lo < m <=
// MC = CT NT / QG candidates (MC = MCAP takes any larger ones, capped).
template <int D, int NT, int CT, int QG, int PQ, typename T>
__device__ __forceinline__ void id_up_node_tile(const H2Dev &h, const int node, int m_lo, double tol)
{
    constexpr int PMAX = PQ, PSTRIDE = H2Cfg<D>::PMAX, RMAX = H2Cfg<D>::RMAX, NCH = H2Cfg<D>::NCH;
    constexpr int QT = PMAX / QG, NGRP = NT / QG, MC = NGRP * CT, NW = NT / 64;
    static_assert(QT * QG == PMAX && QT % 4 == 0 && MC <= NT && NT >= PMAX && NT % 64 == 0 &&
                  PMAX <= PSTRIDE, "tile shape");
    static_assert(MC <= 511, "pivot keys carry 9-bit candidate indices");
    __shared__ float4 s_xin[MC];
    __shared__ double s_cy[MC][D];
    __shared__ double s_prox[PMAX][D];
    __shared__ double s_part[NW][PMAX];
    __shared__ double s_scale[PMAX];
    __shared__ T s_Q[2][PMAX];
    __shared__ float s_U[MC][RMAX];
    __shared__ double s_R[RMAX][RMAX];
    __shared__ int s_piv[RMAX];
    __shared__ unsigned s_key[2][NW];
    __shared__ T s_bv;
    __shared__ int s_base;

    const int t = threadIdx.x;
    const int lane = t & 63, wid = t >> 6;
    const int g = t / QG, qg = t % QG;
    const int ch0 = h.child0[node];
    const bool leafn = ch0 < 0;
    const int l = h.level[node];
    const float4 c = h.cent[node];
    const RootInfo root = *h.root;
    const int kq = min(max(root.k0 - l, 0), kProxyLevels - 1);
    const int p = min(h.proxy_cnt[kq], PMAX);
    double px[3] = {0.0, 0.0, 0.0};
    if (t < p)
        for (int d = 0; d < D; d++) px[d] = h.proxy_pts[((size_t)kq * PSTRIDE + t) * D + d];
    int nb = 0, mt = 0, nc = 0;
    int rk[NCH];
    if (leafn) {
        nb = h.nb[node];
        mt = h.ne[node] - nb;
    } else {
        nc = h.nchild[node];
#pragma unroll
        for (int q = 0; q < NCH; q++) rk[q] = (q < nc) ? h.rank[ch0 + q] : 0;
#pragma unroll
        for (int q = 0; q < NCH; q++) mt += rk[q];
    }
    if (mt <= m_lo || (mt > MC && MC < H2Cfg<D>::MCAP)) return;
    const int m = min(mt, MC);
    // U region of the node (as id_up_node): R entries (FP64) during the ID, then U
    const int part = blockIdx.x % kPoolParts;
    const long long part_cap = pool_part_cap(h, RMAX);
    if (t == 0)
        s_base = leafn ? nb * 2 * RMAX
                       : (int)(2LL * h.n * RMAX + part * part_cap +
                               atomicAdd(&h.cnt[CNT_POOL + part * kCntSlot], 2 * m * RMAX));
    const double cd[3] = {(double)c.x, (double)c.y, (double)c.z};

    // ---- candidate t: coordinates and its row of X
    double my[3] = {0.0, 0.0, 0.0};
    float4 xin = make_float4(0.f, 0.f, 0.f, 0.f);
    if (leafn) {
        if (t < m) {
            const VT<D> y = ((const VT<D> *)h.Ys)[nb + t];
            my[0] = y.x;
            my[1] = y.y;
            xin = make_float4(1.f, y.x - c.x, y.y - c.y, 0.f);
            if constexpr (D == 3) {
                my[2] = y.z;
                xin.w = y.z - c.z;
            }
        }
    } else {
        int cj = -1, k = 0, acc = 0;
#pragma unroll
        for (int q = 0; q < NCH; q++) {
            if (t >= acc && t < acc + rk[q]) {
                cj = ch0 + q;
                k = t - acc;
            }
            acc += rk[q];
        }
        if (cj >= 0) {
            for (int d = 0; d < D; d++) my[d] = h.skel[((size_t)cj * RMAX + k) * D + d];
            const float4 w = h.wc[(size_t)cj * RMAX + k];
            const float4 cc = h.cent[cj];
            xin = make_float4(w.x, w.y + (cc.x - c.x) * w.x, w.z + (cc.y - c.y) * w.x,
                              w.w + (cc.z - c.z) * w.x);
        }
        if (nc == 1) {
            // single child: same skeleton, U = I, coefficients moved to this centre
            __syncthreads();
            const int base = s_base;
            const bool fits = (long long)base + 2LL * m * RMAX <=
                              2LL * h.n * RMAX + (part + 1) * part_cap;
            if (t == 0) {
                if (!fits) atomicOr(&h.flags[FLAG_POOL_OVF], 1);
                h.rank[node] = m;
                h.uoff[node] = base;
            }
            if (t < m) {
                if (fits)
                    for (int q = 0; q < m; q++) h.upool[(size_t)base + (size_t)t * m + q] = (q == t) ? 1.f : 0.f;
                for (int d = 0; d < D; d++) h.skel[((size_t)node * RMAX + t) * D + d] = my[d];
                h.skelf[(size_t)node * RMAX + t] =
                    make_float4((float)my[0], (float)my[1], (float)my[2], 0.f);
                h.wc[(size_t)node * RMAX + t] = xin;
            }
            return;
        }
    }
    if (t < MC) {
        s_xin[t] = xin;
        for (int d = 0; d < D; d++) s_cy[t][d] = my[d];
    }
    if (t < PMAX)
        for (int d = 0; d < D; d++) s_prox[t][d] = cd[d] + px[d];
    for (int e = t; e < RMAX * RMAX; e += NT) s_R[e / RMAX][e % RMAX] = (e / RMAX == e % RMAX) ? 1.0 : 0.0;
    __syncthreads();

    // ---- this thread's tile of A, then column norms over the candidates
    T a[CT][QT];
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int ci = g * CT + ct;
#pragma unroll
        for (int j = 0; j < QT; j++)
            a[ct][j] = (ci < m) ? (T)kern64(s_cy[ci], s_prox[qg * QT + j], D) : T(0);
    }
    {
        // a column's sum of squares: over the thread's CT candidates, then over the groups
        // of the wavefront (lanes ^QG, ^2QG, ..., ^32), then over the wavefronts
        double cs[QT];
#pragma unroll
        for (int j = 0; j < QT; j++) {
            double s = 0.0;
#pragma unroll
            for (int ct = 0; ct < CT; ct++) s = fma((double)a[ct][j], (double)a[ct][j], s);
            cs[j] = s;
        }
#pragma unroll
        for (int j = 0; j < QT; j++)
#pragma unroll
            for (int o = QG; o < 64; o <<= 1) cs[j] += __shfl_xor(cs[j], o, 64);
        if (lane < QG)
#pragma unroll
            for (int j = 0; j < QT; j++) s_part[wid][qg * QT + j] = cs[j];
    }
    __syncthreads();
    for (int q = t; q < PMAX; q += NT) {
        double s2 = 0.0;
        for (int w = 0; w < NW; w++) s2 += s_part[w][q];
        s_scale[q] = (q < p && s2 > 0.0) ? 1.0 / sqrt(s2) : 0.0;
    }
    __syncthreads();
    T nrm[CT];
    unsigned live = 0;  // bit ct: candidate g CT + ct exists and is not yet a pivot
    {
        T sc[QT];
#pragma unroll
        for (int j = 0; j < QT; j++) sc[j] = (T)s_scale[qg * QT + j];
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            T n0 = 0, n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
            for (int j = 0; j < QT; j += 4) {
                a[ct][j] *= sc[j];
                a[ct][j + 1] *= sc[j + 1];
                a[ct][j + 2] *= sc[j + 2];
                a[ct][j + 3] *= sc[j + 3];
                n0 = fma(a[ct][j], a[ct][j], n0);
                n1 = fma(a[ct][j + 1], a[ct][j + 1], n1);
                n2 = fma(a[ct][j + 2], a[ct][j + 2], n2);
                n3 = fma(a[ct][j + 3], a[ct][j + 3], n3);
            }
            nrm[ct] = sum_group<QG>((n0 + n1) + (n2 + n3));
            if (g * CT + ct < m) live |= 1u << ct;
        }
    }

    // ---- column-pivoted MGS; R[k][i] of candidate i goes to rpool[k * m + i] (written by
    // lane 0 of the candidate's group)
    const int rbase = s_base;
    const bool rfits = leafn || (long long)rbase + 2LL * m * RMAX <= 2LL * h.n * RMAX + (part + 1) * part_cap;
    double *rpool = reinterpret_cast<double *>(h.upool + rbase);
    T nrm0 = 0;
    int r = 0;
    const int rlim = min(RMAX, min(m, p));
    const T tol2 = (T)(tol * tol);
    for (int k = 0; k < RMAX; k++) {
        unsigned key = 0;
#pragma unroll
        for (int ct = 0; ct < CT; ct++)
            key = max(key, pivot_key((double)nrm[ct], g * CT + ct, (live >> ct) & 1u));
        key = wave_max_u32(key);
        if (NW > 1) {
            if (lane == 0) s_key[k & 1][wid] = key;
            __syncthreads();
#pragma unroll
            for (int w = 0; w < NW; w++) key = max(key, s_key[k & 1][w]);
        }
        const int bi = key ? 511 - (int)(key & 0x1FFu) : -1;
        const int gb = bi >= 0 ? bi / CT : -1;
        const int cb = bi - gb * CT;
        // the pivot's group publishes its exact norm and, speculatively, its Q row
        if (g == gb) {
            T nb_ = nrm[0];
#pragma unroll
            for (int ct = 1; ct < CT; ct++)
                if (ct == cb) nb_ = nrm[ct];
            const T inv = (T)(1.0 / sqrt((double)nb_));
#pragma unroll
            for (int j = 0; j < QT; j++) {
                T v = a[0][j];
#pragma unroll
                for (int ct = 1; ct < CT; ct++)
                    if (ct == cb) v = a[ct][j];
                s_Q[k & 1][qg * QT + j] = v * inv;
            }
            if (qg == 0) {
                s_bv = nb_;
                s_piv[k] = bi;
            }
        }
        __syncthreads();
        const T bv = key ? s_bv : T(0);
        if (k == 0) nrm0 = bv;
        if (k >= rlim || !(bv > tol2 * nrm0) || !(bv > T(0))) break;
        if (g == gb) {
            live &= ~(1u << cb);
            if (qg == 0) s_R[k][k] = sqrt((double)bv);
        }
        if (live) {  // the same in the group's eight lanes
            T qv[QT];
#pragma unroll
            for (int j = 0; j < QT; j++) qv[j] = s_Q[k & 1][qg * QT + j];
            T dot[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                T d0 = 0, d1 = 0, d2 = 0, d3 = 0;
#pragma unroll
                for (int j = 0; j < QT; j += 4) {
                    d0 = fma(qv[j], a[ct][j], d0);
                    d1 = fma(qv[j + 1], a[ct][j + 1], d1);
                    d2 = fma(qv[j + 2], a[ct][j + 2], d2);
                    d3 = fma(qv[j + 3], a[ct][j + 3], d3);
                }
                dot[ct] = sum_group<QG>((d0 + d1) + (d2 + d3));  // R[k][g CT + ct]
            }
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                if (!((live >> ct) & 1u)) continue;
                if (qg == 0 && rfits) rpool[(size_t)k * m + g * CT + ct] = (double)dot[ct];
                if constexpr (kDowndate && D == 3 && std::is_same<T, double>::value) {
                    // FP64: the residual norm downdated, |a - dot q|^2 = |a|^2 - dot^2 (q
                    // of unit norm), as LAPACK's pivoted QR does.  The pivoting stops at a
                    // residual ratio of tol^2 >= 1e-6, where the downdate's rounding
                    // (~1e-16 of the first norm per step) is 1e-9 of the norms compared.
#pragma unroll
                    for (int j = 0; j < QT; j++) a[ct][j] = fma(-dot[ct], qv[j], a[ct][j]);
                    nrm[ct] = fma(-dot[ct], dot[ct], nrm[ct]);
                } else {
                    T n0 = 0, n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
                    for (int j = 0; j < QT; j += 4) {
                        a[ct][j] = fma(-dot[ct], qv[j], a[ct][j]);
                        a[ct][j + 1] = fma(-dot[ct], qv[j + 1], a[ct][j + 1]);
                        a[ct][j + 2] = fma(-dot[ct], qv[j + 2], a[ct][j + 2]);
                        a[ct][j + 3] = fma(-dot[ct], qv[j + 3], a[ct][j + 3]);
                        n0 = fma(a[ct][j], a[ct][j], n0);
                        n1 = fma(a[ct][j + 1], a[ct][j + 1], n1);
                        n2 = fma(a[ct][j + 2], a[ct][j + 2], n2);
                        n3 = fma(a[ct][j + 3], a[ct][j + 3], n3);
                    }
                    nrm[ct] = sum_group<QG>((n0 + n1) + (n2 + n3));
                }
            }
        }
        r = k + 1;
        if (k == RMAX - 1 && t == 0) atomicOr(&h.flags[FLAG_RANK_CAP], 1);
    }
    __syncthreads();  // every R entry of the pool written

    // ---- one candidate per thread: its column of R, the solve, U
    int kk = -1;
    for (int k = 0; k < r; k++)
        if (s_piv[k] == t) kk = k;
    auto tail = [&](auto rb) {
        constexpr int RB = decltype(rb)::value;
        double rt[RB];
        const int rows = (kk >= 0) ? kk : r;
#pragma unroll
        for (int k = 0; k < RB; k++)
            rt[k] = (t < m && k < rows && rfits) ? rpool[(size_t)k * m + t] : 0.0;
        if (kk >= 0) {
#pragma unroll
            for (int j = 0; j < RB; j++)
                if (j < kk) s_R[j][kk] = rt[j];
        }
        __syncthreads();
        if (t < RB) s_part[0][t] = 1.0 / s_R[t][t];
        __syncthreads();
        if (kk >= 0) {
#pragma unroll
            for (int k = 0; k < RB; k++) rt[k] = (k == kk) ? 1.0 : 0.0;
        } else if (t < m) {
#pragma unroll
            for (int k = RB - 1; k >= 0; k--) {
                double v = rt[k];
#pragma unroll
                for (int j = k + 1; j < RB; j++) v = fma(-s_R[k][j], rt[j], v);
                rt[k] = v * s_part[0][k];
            }
        }
        if (t < MC)
#pragma unroll
            for (int k = 0; k < RB; k++) s_U[t][k] = (float)rt[k];
    };
    if (r <= RMAX / 4)
        tail(std::integral_constant<int, RMAX / 4>{});
    else if (r <= RMAX / 2)
        tail(std::integral_constant<int, RMAX / 2>{});
    else
        tail(std::integral_constant<int, RMAX>{});

    // ---- store U (pool), skeleton, and the upward coefficients
    if (t == 0) {
        if (!rfits) atomicOr(&h.flags[FLAG_POOL_OVF], 1);
        h.rank[node] = r;
        h.uoff[node] = s_base;
    }
    if (kk >= 0) {
        for (int d = 0; d < D; d++) h.skel[((size_t)node * RMAX + kk) * D + d] = my[d];
        h.skelf[(size_t)node * RMAX + kk] = make_float4((float)my[0], (float)my[1], (float)my[2], 0.f);
    }
    __syncthreads();
    const int base = s_base;
    if (rfits)
        for (int e = t; e < m * r; e += NT) h.upool[(size_t)base + e] = s_U[e / r][e % r];
    if (t < r) {
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int u = 0; u < m; u++) {
            const float uk = s_U[u][t];
            const float4 x = s_xin[u];
            w.x = fmaf(uk, x.x, w.x);
            w.y = fmaf(uk, x.y, w.y);
            w.z = fmaf(uk, x.z, w.z);
            w.w = fmaf(uk, x.w, w.w);
        }
        h.wc[(size_t)node * RMAX + t] = w;
    }
}

template <int D, int NT, int CT, int QG, int PQ, bool STRIDE, typename T>
__global__ void __launch_bounds__(NT)
id_tile_kernel(H2Dev h, const int *__restrict__ list, int lb, int count, int m_lo, double tol)
{
    if constexpr (STRIDE) {
        for (int b = blockIdx.x; b < count; b += gridDim.x) {
            const int node = list ? list[b] : lb + b;
            __syncthreads();  // the previous node's LDS reads are complete
            if (!list && h.child0[node] < 0) continue;
            id_up_node_tile<D, NT, CT, QG, PQ, T>(h, node, m_lo, tol);
        }
    } else {
        const int node = list ? list[blockIdx.x] : lb + blockIdx.x;
        if (!list && h.child0[node] < 0) return;
        id_up_node_tile<D, NT, CT, QG, PQ, T>(h, node, m_lo, tol);
    }
}

// list != nullptr: the nodes list[0..count) (the leaves); otherwise the internal nodes
// among [lb, lb + count).  One workgroup per node at a time, in one size class.  STRIDE:
// a grid smaller than `count` walks the nodes with a stride, so a launch whose class
// holds few of them costs few workgroups; otherwise one workgroup per node (the grid is
// `count`, and the loop's state does not compete with the ID for registers).
template <int D, int NT, int PQ, bool STRIDE, typename T>
__global__ void __launch_bounds__(NT)
id_up_kernel(H2Dev h, const int *__restrict__ list, int lb, int count, int m_lo, double tol)
{
    if constexpr (STRIDE) {
        for (int b = blockIdx.x; b < count; b += gridDim.x) {
            const int node = list ? list[b] : lb + b;
            __syncthreads();  // the previous node's LDS reads are complete
            if (!list && h.child0[node] < 0) continue;  // leaves were done in their own launch
            id_up_node<D, NT, PQ, T>(h, node, m_lo, tol);
        }
    } else {
        const int node = list ? list[blockIdx.x] : lb + blockIdx.x;
        if (!list && h.child0[node] < 0) return;
        id_up_node<D, NT, PQ, T>(h, node, m_lo, tol);
    }
}

// Nodes with at most 64 candidates run in single-wavefront workgroups (their barriers
// are free), larger ones in 128 threads (2D).  A level of at most kOneClassLevel nodes
// (the top of the tree, where every launch is one node's latency) runs all its nodes in
// one launch of the largest class instead of one per class.
// 3D (up to 8 x 48 candidates, 96 proxies): the 384-thread class needs ~106 KB of LDS, so
// one workgroup per CU with six waves behind every barrier.  Leaves (tens of points
// typically) run in classes of 64 and 128 threads (~60 KB, two per CU), on grids that
// walk the leaf list with a stride so that a class's launch costs few workgroups when it
// holds few leaves.  Internal levels of at most kOneClassLevel3 nodes run in one launch of the
// 384-thread class; larger ones in the 64, 128 and 384 classes with one workgroup per node
// (walking single-child nodes one after another in a strided grid cost more than the
// smaller classes saved).
constexpr int kOneClassLevel = 512;
constexpr int kOneClassLevel3 = 2048;  // 3D internal levels: one 384-thread launch up to this

// kOneClassLevel3, or TSNE_ID_ONECLASS3 when set (read once)
int one_class_level3()
{
    static const int v = [] {
        const char *e = getenv("TSNE_ID_ONECLASS3");
        const int x = e ? atoi(e) : kOneClassLevel3;
        return x >= 0 ? x : kOneClassLevel3;
    }();
    return v;
}
// IDs in tiles (id_tile_kernel), unless TSNE_ID_TILE3=0 (3D) / TSNE_ID_TILE2=0 (2D) select
// the one-thread-per-candidate kernel (read once)
bool tile3()
{
    static const bool v = [] {
        const char *e = getenv("TSNE_ID_TILE3");
        return !(e && e[0] == '0');
    }();
    return v;
}
bool tile2()
{
    static const bool v = [] {
        const char *e = getenv("TSNE_ID_TILE2");
        return !(e && e[0] == '0');
    }();
    return v;
}
constexpr int kLeafGridSmall = 8192;  // workgroups of a 3D leaf launch, 64/128 threads
constexpr int kLeafGridLarge = 1024;  // same, 384 threads (leaves of > 128 points)

template <typename T>
void launch_id_up_t(int D, const H2Dev &h, const int *list, int lb, int count, double tol, int pq,
                    int one_class, hipStream_t st)
{
    const int oc2 = one_class >= 0 ? one_class : kOneClassLevel;
    const int oc3 = one_class >= 0 ? one_class : one_class_level3();
    if (D == 2 && tile2()) {
        // 2D in tiles of 2 candidates x PQ/4 proxy slots: classes of 64 (128 threads) and
        // 128 candidates (256 threads)
        const bool one = count <= oc2;
        if (pq <= 32) {
            if (!one)
                hipLaunchKernelGGL((id_tile_kernel<2, 128, 2, 4, 32, false, T>), dim3(count), dim3(128), 0,
                                   st, h, list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_tile_kernel<2, 256, 2, 4, 32, false, T>), dim3(count), dim3(256), 0, st,
                               h, list, lb, count, one ? 0 : 64, tol);
        } else {
            if (!one)
                hipLaunchKernelGGL((id_tile_kernel<2, 128, 2, 4, 48, false, T>), dim3(count), dim3(128), 0,
                                   st, h, list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_tile_kernel<2, 256, 2, 4, 48, false, T>), dim3(count), dim3(256), 0, st,
                               h, list, lb, count, one ? 0 : 64, tol);
        }
    } else if (D == 2) {
        const bool one = count <= oc2;
        if (pq <= 32) {
            if (!one)
                hipLaunchKernelGGL((id_up_kernel<2, 64, 32, false, T>), dim3(count), dim3(64), 0, st, h,
                                   list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_up_kernel<2, 128, 32, false, T>), dim3(count), dim3(128), 0, st, h,
                               list, lb, count, one ? 0 : 64, tol);
        } else {
            if (!one)
                hipLaunchKernelGGL((id_up_kernel<2, 64, 48, false, T>), dim3(count), dim3(64), 0, st, h,
                                   list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_up_kernel<2, 128, 48, false, T>), dim3(count), dim3(128), 0, st, h,
                               list, lb, count, one ? 0 : 64, tol);
        }
    } else if (tile3()) {
        // 3D in tiles: classes of 64 (128 threads), 128 (256) and 384 candidates (768)
        if (!list && count <= oc3) {
            hipLaunchKernelGGL((id_tile_kernel<3, 768, 4, 8, 96, false, T>), dim3(count), dim3(768), 0, st, h,
                               list, lb, count, 0, tol);
        } else if (!list) {
            hipLaunchKernelGGL((id_tile_kernel<3, 128, 4, 8, 96, false, T>), dim3(count), dim3(128), 0, st, h,
                               list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_tile_kernel<3, 256, 4, 8, 96, false, T>), dim3(count), dim3(256), 0, st, h,
                               list, lb, count, 64, tol);
            hipLaunchKernelGGL((id_tile_kernel<3, 768, 4, 8, 96, false, T>), dim3(count), dim3(768), 0, st, h,
                               list, lb, count, 128, tol);
        } else {
            const int gs = std::min(count, kLeafGridSmall), gl = std::min(count, kLeafGridLarge);
            hipLaunchKernelGGL((id_tile_kernel<3, 128, 4, 8, 96, true, T>), dim3(gs), dim3(128), 0, st, h,
                               list, lb, count, 0, tol);
            hipLaunchKernelGGL((id_tile_kernel<3, 256, 4, 8, 96, true, T>), dim3(gs), dim3(256), 0, st, h,
                               list, lb, count, 64, tol);
            hipLaunchKernelGGL((id_tile_kernel<3, 768, 4, 8, 96, true, T>), dim3(gl), dim3(768), 0, st, h,
                               list, lb, count, 128, tol);
        }
    } else if (!list && count <= oc3) {
        hipLaunchKernelGGL((id_up_kernel<3, 384, 96, false, T>), dim3(count), dim3(384), 0, st, h, list,
                           lb, count, 0, tol);
    } else if (!list) {
        // a large internal level: one workgroup per node in each class (no stride: many of
        // these nodes take the single-child shortcut, and the small classes' early exits
        // are cheap), so that nodes of <= 128 candidates run at two to four per CU
        hipLaunchKernelGGL((id_up_kernel<3, 64, 96, false, T>), dim3(count), dim3(64), 0, st, h, list,
                           lb, count, 0, tol);
        hipLaunchKernelGGL((id_up_kernel<3, 128, 96, false, T>), dim3(count), dim3(128), 0, st, h, list,
                           lb, count, 64, tol);
        hipLaunchKernelGGL((id_up_kernel<3, 384, 96, false, T>), dim3(count), dim3(384), 0, st, h, list,
                           lb, count, 128, tol);
    } else {
        const int gs = std::min(count, kLeafGridSmall), gl = std::min(count, kLeafGridLarge);
        hipLaunchKernelGGL((id_up_kernel<3, 64, 96, true, T>), dim3(gs), dim3(64), 0, st, h, list, lb,
                           count, 0, tol);
        hipLaunchKernelGGL((id_up_kernel<3, 128, 96, true, T>), dim3(gs), dim3(128), 0, st, h, list, lb,
                           count, 64, tol);
        hipLaunchKernelGGL((id_up_kernel<3, 384, 96, true, T>), dim3(gl), dim3(384), 0, st, h, list, lb,
                           count, 128, tol);
    }
}

int launch_id_up(int D, const H2Dev &h, const int *list, int lb, int count, double tol, int pq,
                 bool f32, int one_class, hipStream_t st)
{
    if (count <= 0) return 0;
    if (f32)
        launch_id_up_t<float>(D, h, list, lb, count, tol, pq, one_class, st);
    else
        launch_id_up_t<double>(D, h, list, lb, count, tol, pq, one_class, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsne
