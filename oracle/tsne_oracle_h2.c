/*
 * tsne_oracle_h2.c -- CPU (OpenMP, FP64) restatement of the reference's H2 repulsive
 * product and its rebuild.  TEST INFRASTRUCTURE ONLY, like tsne_oracle.c: it is loaded by
 * tests/ and by bench.py's cpu_baseline leg (the reference's algorithm timed on the host
 * cores); the product library never links it.
 *
 * The reference builds an H2 matrix of K(x, y) = (1 + |x - y|^2)^-2 with H2Pack every
 * h2_rebuild_freq iterations (src/tsne_h2pack.c:772-953: proxy-point interpolative
 * decomposition, QR_REL_NRM, leaf size h2_leaf_size, ID tolerance from h2_tol) and takes
 * s = K 1 and v = K Y from one fused product (src/tsne_h2_matmul.c:225-452;
 * src/tsne_h2pack.c:538-546); H2Pack itself is not part of the reference's sources, so
 * this file restates the published method:
 *   tree      cube around the embedding, points sorted by Morton key (LSD radix sort),
 *             nodes split by key digit while they hold more than `leaf` points;
 *   lists     dual-tree traversal from (root, root): a pair of boxes whose L-infinity gap
 *             is at least twice the larger half-side is admissible (each lies outside the
 *             other's proxy region, the cube of three half-sides), two leaves otherwise
 *             interact directly, else the larger node is split;
 *   proxies   per level, points on square (cube) shells at L-infinity radii 3h x {1, 1.3,
 *             1.8, 2.6, 4, 7, 15} around the node centre (h the level's half-side);
 *   ID        rows of A = K(candidates, proxies), columns normalised, by column-pivoted
 *             modified Gram-Schmidt on A^T until the largest residual norm falls below
 *             tol times the first; U = [I; R11^-1 R12]^T on the candidates; candidates are
 *             a leaf's points or the children's skeleton points (nested bases);
 *   product   upward W = U^T [1, y] (nested), coupling over admissible pairs, downward
 *             through the transfer blocks, near field exactly; s and v from [1, y] columns.
 * Pinned (tests/test_oracle_h2.py) against oracle_repulsive_exact: relative error of the
 * repulsive force and of Z within h2_tol on t-SNE embeddings in 2D and 3D, and full-run
 * KL within 1% of the exact-repulsion run on the same schedule.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double oracle_repulsive_exact(int n, int dim, const double *Y, double *rep, double *s_out,
                              double *v_out);
void oracle_attractive(int n, int dim, const int *row_ptr, const int *col_idx,
                       const double *val, const double *Y, double exag, double *grad);
double oracle_kl(int n, int dim, const int *row_ptr, const int *col_idx, const double *val,
                 const double *Y, double Z, double exag);
void oracle_update(int n, int dim, double *Y, double *uY, double *gains, double *grad,
                   double momentum, double lr);
int oracle_is_rebuild_iter(int it, int freq, int freq_early, int early_exag_iter, int reuse);

#define OH2_MAXLEV 32

typedef struct {
    int n, dim, leaf, nc;      /* nc = 2^dim children per split */
    double tol;
    double c0[3], H0;          /* root cube */
    int *perm;                 /* sorted position k -> caller index */
    double *Ys;                /* sorted coordinates, n x dim row-major */
    /* nodes in breadth-first order (children of a node contiguous) */
    int nn, ncap;
    int *nb, *ne, *lev, *ch0, *nch, *par;
    double *cen;               /* nn x 3 */
    int lvl[OH2_MAXLEV + 2];   /* level l is nodes [lvl[l], lvl[l+1]) */
    int nlev;
    /* lists, CSR by target node */
    int *far_ptr, *far_idx, *near_ptr, *near_idx;
    long long n_far, n_near;
    /* ID: rank, U (m x r, row-major) and skeleton coordinates (r x dim) per node */
    int *m, *rank;
    double **U, **sk;
    /* proxies per level (relative to the centre), count per level */
    double *prox[OH2_MAXLEV + 1];
    int np[OH2_MAXLEV + 1];
} oh2_t;

static double kern(const double *a, const double *b, int dim)
{
    double r2 = 1.0;
    for (int d = 0; d < dim; d++) {
        const double t = a[d] - b[d];
        r2 += t * t;
    }
    return 1.0 / (r2 * r2);
}

/* ---------------------------------------------------------------------- sort */

static uint64_t spread2(uint64_t x)   /* 30 bits -> every second bit */
{
    x &= 0x3fffffffULL;
    x = (x | (x << 16)) & 0x0000ffff0000ffffULL;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffULL;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | (x << 2)) & 0x3333333333333333ULL;
    x = (x | (x << 1)) & 0x5555555555555555ULL;
    return x;
}

static uint64_t spread3(uint64_t x)   /* 21 bits -> every third bit */
{
    x &= 0x1fffffULL;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

/* Stable LSD radix sort of (key, index) by key, 8-bit digits, `bits` key bits. */
static void radix_sort(uint64_t *key, int *idx, int n, int bits)
{
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    int *i2 = (int *)malloc(sizeof(int) * (size_t)n);
    int ntmax = 1;
#ifdef _OPENMP
    ntmax = omp_get_max_threads();
#endif
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * 256 * (size_t)ntmax);
    for (int sh = 0; sh < bits; sh += 8) {
        memset(cnt, 0, sizeof(size_t) * 256 * (size_t)ntmax);
        #pragma omp parallel num_threads(ntmax)
        {
            int t = 0, nt = 1;  /* the team actually granted */
#ifdef _OPENMP
            t = omp_get_thread_num();
            nt = omp_get_num_threads();
#endif
            const int b = (int)((long long)n * t / nt), e = (int)((long long)n * (t + 1) / nt);
            size_t *c = cnt + 256 * (size_t)t;
            for (int i = b; i < e; i++) c[(key[i] >> sh) & 255]++;
            #pragma omp barrier
            #pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 256; d++)
                    for (int u = 0; u < nt; u++) {
                        const size_t x = cnt[256 * (size_t)u + d];
                        cnt[256 * (size_t)u + d] = run;
                        run += x;
                    }
            }
            for (int i = b; i < e; i++) {
                const size_t o = c[(key[i] >> sh) & 255]++;
                k2[o] = key[i];
                i2[o] = idx[i];
            }
        }
        memcpy(key, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(idx, i2, sizeof(int) * (size_t)n);
    }
    free(k2); free(i2); free(cnt);
}

/* ---------------------------------------------------------------------- tree */

static int add_node(oh2_t *h, int nb, int ne, int lev, int par, const double *cen)
{
    if (h->nn == h->ncap) {
        h->ncap *= 2;
        h->nb = (int *)realloc(h->nb, sizeof(int) * (size_t)h->ncap);
        h->ne = (int *)realloc(h->ne, sizeof(int) * (size_t)h->ncap);
        h->lev = (int *)realloc(h->lev, sizeof(int) * (size_t)h->ncap);
        h->ch0 = (int *)realloc(h->ch0, sizeof(int) * (size_t)h->ncap);
        h->nch = (int *)realloc(h->nch, sizeof(int) * (size_t)h->ncap);
        h->par = (int *)realloc(h->par, sizeof(int) * (size_t)h->ncap);
        h->cen = (double *)realloc(h->cen, sizeof(double) * 3 * (size_t)h->ncap);
    }
    const int i = h->nn++;
    h->nb[i] = nb; h->ne[i] = ne; h->lev[i] = lev; h->par[i] = par;
    h->ch0[i] = -1; h->nch[i] = 0;
    for (int d = 0; d < 3; d++) h->cen[3 * (size_t)i + d] = cen[d];
    return i;
}

static double half_side(const oh2_t *h, int l) { return h->H0 / (double)(1ULL << l); }

/* ---------------------------------------------------------------------- lists */

typedef struct { int *a; long long n, cap; } ivec;

static void ipush2(ivec *v, int x, int y)
{
    if (v->n + 2 > v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->a = (int *)realloc(v->a, sizeof(int) * (size_t)v->cap);
    }
    v->a[v->n++] = x;
    v->a[v->n++] = y;
}

static int admissible(const oh2_t *h, int i, int j)
{
    const double hi = half_side(h, h->lev[i]), hj = half_side(h, h->lev[j]);
    const double hm = hi > hj ? hi : hj;
    double gap = -1e300;
    for (int d = 0; d < h->dim; d++) {
        const double g = fabs(h->cen[3 * (size_t)i + d] - h->cen[3 * (size_t)j + d]) - hi - hj;
        if (g > gap) gap = g;
    }
    return gap >= 2.0 * hm * (1.0 - 1e-12);
}

static void interact(const oh2_t *h, int i, int j, ivec *far, ivec *near)
{
    if (i != j && admissible(h, i, j)) {
        ipush2(far, i, j);
        return;
    }
    const int li = h->nch[i] == 0, lj = h->nch[j] == 0;
    if (li && lj) {
        ipush2(near, i, j);
        return;
    }
    if (li || (!lj && h->lev[j] < h->lev[i])) {
        for (int c = 0; c < h->nch[j]; c++) interact(h, i, h->ch0[j] + c, far, near);
    } else {
        for (int c = 0; c < h->nch[i]; c++) interact(h, h->ch0[i] + c, j, far, near);
    }
}

static void to_csr(const ivec *v, int nn, int **ptr, int **idx)
{
    const long long np = v->n / 2;
    *ptr = (int *)calloc((size_t)nn + 1, sizeof(int));
    *idx = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    for (long long k = 0; k < np; k++) (*ptr)[v->a[2 * k] + 1]++;
    for (int i = 0; i < nn; i++) (*ptr)[i + 1] += (*ptr)[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)nn);
    memcpy(fill, *ptr, sizeof(int) * (size_t)nn);
    for (long long k = 0; k < np; k++) (*idx)[fill[v->a[2 * k]]++] = v->a[2 * k + 1];
    free(fill);
}

/* ---------------------------------------------------------------------- proxies */

static void make_proxies(oh2_t *h, int l)
{
    static const double f[] = {1.0, 1.3, 1.8, 2.6, 4.0, 7.0, 15.0};
    const int nsh = (int)(sizeof f / sizeof f[0]);
    const double hs = half_side(h, l);
    const int q = h->dim == 2 ? 4 : 3;   /* 2D: 4 points per side; 3D: a 3 x 3 grid per face */
    const int per = h->dim == 2 ? 4 * q : 6 * q * q;
    h->np[l] = nsh * per;
    h->prox[l] = (double *)malloc(sizeof(double) * (size_t)h->np[l] * h->dim);
    int k = 0;
    for (int s = 0; s < nsh; s++) {
        const double r = 3.0 * hs * f[s];
        if (h->dim == 2) {
            for (int side = 0; side < 4; side++)
                for (int t = 0; t < q; t++) {
                    const double u = -r + 2.0 * r * (t + 0.5 * (s & 1)) / q;  /* shells staggered */
                    double x = 0, y = 0;
                    if (side == 0) { x = u; y = -r; }
                    else if (side == 1) { x = r; y = u; }
                    else if (side == 2) { x = -u; y = r; }
                    else { x = -r; y = -u; }
                    h->prox[l][2 * k] = x;
                    h->prox[l][2 * k + 1] = y;
                    k++;
                }
        } else {
            for (int face = 0; face < 6; face++)
                for (int a = 0; a < q; a++)
                    for (int b = 0; b < q; b++) {
                        const double u = -r + 2.0 * r * (a + 0.5) / q, v = -r + 2.0 * r * (b + 0.5) / q;
                        const int ax = face >> 1;
                        const double w = (face & 1) ? r : -r;
                        double p[3];
                        p[ax] = w;
                        p[(ax + 1) % 3] = u;
                        p[(ax + 2) % 3] = v;
                        for (int d = 0; d < 3; d++) h->prox[l][3 * k + d] = p[d];
                        k++;
                    }
        }
    }
}

/* ---------------------------------------------------------------------- ID */

/* Candidate coordinates of node i (m x dim, row-major) into buf; returns m. */
static int candidates(const oh2_t *h, int i, double *buf)
{
    const int dim = h->dim;
    if (h->nch[i] == 0) {
        const int m = h->ne[i] - h->nb[i];
        memcpy(buf, h->Ys + (size_t)h->nb[i] * dim, sizeof(double) * (size_t)m * dim);
        return m;
    }
    int m = 0;
    for (int c = 0; c < h->nch[i]; c++) {
        const int j = h->ch0[i] + c;
        memcpy(buf + (size_t)m * dim, h->sk[j], sizeof(double) * (size_t)h->rank[j] * dim);
        m += h->rank[j];
    }
    return m;
}

static void node_id(oh2_t *h, int i)
{
    const int dim = h->dim, l = h->lev[i];
    int mc = 0;
    if (h->nch[i] == 0) mc = h->ne[i] - h->nb[i];
    else for (int c = 0; c < h->nch[i]; c++) mc += h->rank[h->ch0[i] + c];
    const int p = h->np[l];
    double *cand = (double *)malloc(sizeof(double) * (size_t)(mc > 0 ? mc : 1) * dim);
    const int m = candidates(h, i, cand);
    h->m[i] = m;
    const double *ci = h->cen + 3 * (size_t)i;
    /* A (m x p): rows are candidates; columns normalised */
    double *A = (double *)malloc(sizeof(double) * (size_t)m * p);
    double pt[3];
    for (int q = 0; q < p; q++) {
        for (int d = 0; d < dim; d++) pt[d] = ci[d] + h->prox[l][(size_t)q * dim + d];
        for (int a = 0; a < m; a++) A[(size_t)a * p + q] = kern(cand + (size_t)a * dim, pt, dim);
    }
    for (int q = 0; q < p; q++) {
        double s2 = 0.0;
        for (int a = 0; a < m; a++) s2 += A[(size_t)a * p + q] * A[(size_t)a * p + q];
        const double sc = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
        for (int a = 0; a < m; a++) A[(size_t)a * p + q] *= sc;
    }
    /* pivoted MGS over rows; R[k][a] for row a at step k */
    const int rmax = m < p ? m : p;
    double *R = (double *)calloc((size_t)(rmax > 0 ? rmax : 1) * (m > 0 ? m : 1), sizeof(double));
    double *nrm = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    int *piv = (int *)malloc(sizeof(int) * (size_t)(rmax > 0 ? rmax : 1));
    char *sel = (char *)calloc((size_t)(m > 0 ? m : 1), 1);
    double *qv = (double *)malloc(sizeof(double) * (size_t)p);
    for (int a = 0; a < m; a++) {
        double s2 = 0.0;
        for (int q = 0; q < p; q++) s2 += A[(size_t)a * p + q] * A[(size_t)a * p + q];
        nrm[a] = s2;
    }
    double n0 = 0.0;
    int r = 0;
    for (int k = 0; k < rmax; k++) {
        int b = -1;
        double bv = -1.0;
        for (int a = 0; a < m; a++)
            if (!sel[a] && nrm[a] > bv) { bv = nrm[a]; b = a; }
        if (k == 0) n0 = bv;
        if (b < 0 || !(bv > h->tol * h->tol * n0) || !(bv > 0.0)) break;
        sel[b] = 1;
        piv[k] = b;
        const double rk = sqrt(bv);
        R[(size_t)k * m + b] = rk;
        for (int q = 0; q < p; q++) qv[q] = A[(size_t)b * p + q] / rk;
        for (int a = 0; a < m; a++) {
            if (sel[a]) continue;
            double dot = 0.0;
            for (int q = 0; q < p; q++) dot += qv[q] * A[(size_t)a * p + q];
            R[(size_t)k * m + a] = dot;
            double s2 = 0.0;
            for (int q = 0; q < p; q++) {
                A[(size_t)a * p + q] -= dot * qv[q];
                s2 += A[(size_t)a * p + q] * A[(size_t)a * p + q];
            }
            nrm[a] = s2;
        }
        r = k + 1;
    }
    /* U (m x r): pivot rows unit vectors, others R11^-1 R[:, a] */
    h->rank[i] = r;
    h->U[i] = (double *)calloc((size_t)(m > 0 ? m : 1) * (r > 0 ? r : 1), sizeof(double));
    h->sk[i] = (double *)malloc(sizeof(double) * (size_t)(r > 0 ? r : 1) * dim);
    int *slot = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (int a = 0; a < m; a++) slot[a] = -1;
    for (int k = 0; k < r; k++) {
        slot[piv[k]] = k;
        memcpy(h->sk[i] + (size_t)k * dim, cand + (size_t)piv[k] * dim, sizeof(double) * dim);
    }
    double *x = (double *)malloc(sizeof(double) * (size_t)(r > 0 ? r : 1));
    for (int a = 0; a < m; a++) {
        double *u = h->U[i] + (size_t)a * r;
        if (slot[a] >= 0) { u[slot[a]] = 1.0; continue; }
        for (int k = r - 1; k >= 0; k--) {
            double v = R[(size_t)k * m + a];
            for (int j = k + 1; j < r; j++) v -= R[(size_t)k * m + piv[j]] * x[j];
            x[k] = v / R[(size_t)k * m + piv[k]];
        }
        for (int k = 0; k < r; k++) u[k] = x[k];
    }
    free(x); free(slot); free(cand); free(A); free(R); free(nrm); free(piv); free(sel); free(qv);
}

/* ---------------------------------------------------------------------- build */

void oracle_h2_free(void *hv)
{
    oh2_t *h = (oh2_t *)hv;
    if (!h) return;
    free(h->perm); free(h->Ys);
    free(h->nb); free(h->ne); free(h->lev); free(h->ch0); free(h->nch); free(h->par); free(h->cen);
    free(h->far_ptr); free(h->far_idx); free(h->near_ptr); free(h->near_idx);
    if (h->U) for (int i = 0; i < h->nn; i++) { free(h->U[i]); free(h->sk[i]); }
    free(h->U); free(h->sk); free(h->m); free(h->rank);
    for (int l = 0; l <= OH2_MAXLEV; l++) free(h->prox[l]);
    free(h);
}

/* Builds the H2 representation of the embedding Y (dim x n, dimension-major) with leaf
 * size `leaf` and ID tolerance `tol` (the reference's 0.5 h2_tol).  NULL for dim outside
 * {2, 3} or n < 2. */
void *oracle_h2_build(int n, int dim, const double *Y, int leaf, double tol)
{
    if ((dim != 2 && dim != 3) || n < 2 || leaf < 1) return NULL;
    oh2_t *h = (oh2_t *)calloc(1, sizeof(oh2_t));
    h->n = n; h->dim = dim; h->leaf = leaf; h->tol = tol; h->nc = 1 << dim;
    const int B = dim == 2 ? 30 : 21;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        lo[d] = hi[d] = Y[(size_t)d * n];
        for (int i = 1; i < n; i++) {
            const double y = Y[(size_t)d * n + i];
            if (y < lo[d]) lo[d] = y;
            if (y > hi[d]) hi[d] = y;
        }
    }
    h->H0 = 0.0;
    for (int d = 0; d < dim; d++) {
        h->c0[d] = 0.5 * (lo[d] + hi[d]);
        const double e = 0.5 * (hi[d] - lo[d]);
        if (e > h->H0) h->H0 = e;
    }
    h->H0 = h->H0 * (1.0 + 1e-9) + 1e-300;
    /* Morton keys, sorted */
    uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    h->perm = (int *)malloc(sizeof(int) * (size_t)n);
    const double scale = (double)(1ULL << B) / (2.0 * h->H0);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint64_t k = 0;
        for (int d = 0; d < dim; d++) {
            double q = floor((Y[(size_t)d * n + i] - (h->c0[d] - h->H0)) * scale);
            if (q < 0) q = 0;
            if (q > (double)((1ULL << B) - 1)) q = (double)((1ULL << B) - 1);
            const uint64_t u = (uint64_t)q;
            k |= (dim == 2 ? spread2(u) : spread3(u)) << d;
        }
        key[i] = k;
        h->perm[i] = i;
    }
    radix_sort(key, h->perm, n, B * dim);
    h->Ys = (double *)malloc(sizeof(double) * (size_t)n * dim);
    for (int k = 0; k < n; k++)
        for (int d = 0; d < dim; d++) h->Ys[(size_t)k * dim + d] = Y[(size_t)d * n + h->perm[k]];
    /* nodes, breadth first */
    h->ncap = 1024;
    h->nb = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->ne = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->lev = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->ch0 = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->nch = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->par = (int *)malloc(sizeof(int) * (size_t)h->ncap);
    h->cen = (double *)malloc(sizeof(double) * 3 * (size_t)h->ncap);
    double c0[3] = {h->c0[0], h->c0[1], dim == 3 ? h->c0[2] : 0.0};
    add_node(h, 0, n, 0, -1, c0);
    h->lvl[0] = 0;
    int l = 0;
    for (; l < B && l < OH2_MAXLEV; l++) {
        const int b = h->lvl[l], e = h->nn;
        h->lvl[l + 1] = e;
        if (b == e) break;
        const int shift = dim * (B - 1 - l);
        const double hc = half_side(h, l + 1);
        for (int i = b; i < e; i++) {
            if (h->ne[i] - h->nb[i] <= leaf) continue;
            int first = -1, cnt = 0;
            int s = h->nb[i];
            while (s < h->ne[i]) {
                const int dig = (int)((key[s] >> shift) & (uint64_t)(h->nc - 1));
                int t = s + 1;
                while (t < h->ne[i] && (int)((key[t] >> shift) & (uint64_t)(h->nc - 1)) == dig) t++;
                double cc[3] = {0, 0, 0};
                for (int d = 0; d < dim; d++)
                    cc[d] = h->cen[3 * (size_t)i + d] + (((dig >> d) & 1) ? hc : -hc);
                const int c = add_node(h, s, t, l + 1, i, cc);
                if (first < 0) first = c;
                cnt++;
                s = t;
            }
            h->ch0[i] = first;
            h->nch[i] = cnt;
        }
    }
    h->nlev = l + 1;
    if (h->nlev > OH2_MAXLEV + 1) h->nlev = OH2_MAXLEV + 1;
    h->lvl[h->nlev] = h->nn;
    free(key);
    /* lists */
    ivec far = {0}, near = {0};
    interact(h, 0, 0, &far, &near);
    h->n_far = far.n / 2;
    h->n_near = near.n / 2;
    to_csr(&far, h->nn, &h->far_ptr, &h->far_idx);
    to_csr(&near, h->nn, &h->near_ptr, &h->near_idx);
    free(far.a); free(near.a);
    /* proxies and IDs, finest level first (nodes of one level in parallel) */
    for (int q = 1; q < h->nlev; q++) make_proxies(h, q);
    h->m = (int *)calloc((size_t)h->nn, sizeof(int));
    h->rank = (int *)calloc((size_t)h->nn, sizeof(int));
    h->U = (double **)calloc((size_t)h->nn, sizeof(double *));
    h->sk = (double **)calloc((size_t)h->nn, sizeof(double *));
    for (int q = h->nlev - 1; q >= 1; q--) {
        #pragma omp parallel for schedule(dynamic, 4)
        for (int i = h->lvl[q]; i < h->lvl[q + 1]; i++) node_id(h, i);
    }
    return h;
}

/* ---------------------------------------------------------------------- product */

/* Repulsive force of the t-SNE gradient from the H2 product, as oracle_repulsive_exact:
 * s_i = sum_{j != i} K_ij and v_i = sum_{j != i} K_ij y_j from G = K [1, y];
 * Z = sum s + 2 sum |y|^2 s - 2 sum y.v; rep_i = (4/Z)(y_i s_i - v_i) (dim x n).
 * Returns Z. */
double oracle_h2_repulsive(void *hv, const double *Y, double *rep)
{
    oh2_t *h = (oh2_t *)hv;
    const int n = h->n, dim = h->dim, nc = dim + 1, nn = h->nn;
    double **W = (double **)calloc((size_t)nn, sizeof(double *));
    double **Zc = (double **)calloc((size_t)nn, sizeof(double *));
    for (int i = 1; i < nn; i++) {
        W[i] = (double *)calloc((size_t)(h->rank[i] > 0 ? h->rank[i] : 1) * nc, sizeof(double));
        Zc[i] = (double *)calloc((size_t)(h->rank[i] > 0 ? h->rank[i] : 1) * nc, sizeof(double));
    }
    /* upward */
    for (int q = h->nlev - 1; q >= 1; q--) {
        #pragma omp parallel for schedule(dynamic, 8)
        for (int i = h->lvl[q]; i < h->lvl[q + 1]; i++) {
            const int r = h->rank[i], m = h->m[i];
            const double *U = h->U[i];
            double *w = W[i];
            if (h->nch[i] == 0) {
                for (int a = 0; a < m; a++) {
                    const double *y = h->Ys + (size_t)(h->nb[i] + a) * dim;
                    for (int k = 0; k < r; k++) {
                        const double u = U[(size_t)a * r + k];
                        w[(size_t)k * nc] += u;
                        for (int d = 0; d < dim; d++) w[(size_t)k * nc + 1 + d] += u * y[d];
                    }
                }
            } else {
                int a = 0;
                for (int c = 0; c < h->nch[i]; c++) {
                    const int j = h->ch0[i] + c;
                    for (int t = 0; t < h->rank[j]; t++, a++)
                        for (int k = 0; k < r; k++) {
                            const double u = U[(size_t)a * r + k];
                            for (int e = 0; e < nc; e++) w[(size_t)k * nc + e] += u * W[j][(size_t)t * nc + e];
                        }
                }
            }
        }
    }
    /* coupling */
    #pragma omp parallel for schedule(dynamic, 8)
    for (int i = 1; i < nn; i++) {
        const int ri = h->rank[i];
        for (int f = h->far_ptr[i]; f < h->far_ptr[i + 1]; f++) {
            const int j = h->far_idx[f], rj = h->rank[j];
            for (int a = 0; a < ri; a++)
                for (int b = 0; b < rj; b++) {
                    const double k = kern(h->sk[i] + (size_t)a * dim, h->sk[j] + (size_t)b * dim, dim);
                    for (int e = 0; e < nc; e++) Zc[i][(size_t)a * nc + e] += k * W[j][(size_t)b * nc + e];
                }
        }
    }
    /* downward and leaves; near field */
    double *G = (double *)calloc((size_t)n * nc, sizeof(double));
    for (int q = 1; q < h->nlev; q++) {
        #pragma omp parallel for schedule(dynamic, 8)
        for (int i = h->lvl[q]; i < h->lvl[q + 1]; i++) {
            const int r = h->rank[i];
            const double *U = h->U[i];
            if (h->nch[i] == 0) {
                for (int a = 0; a < h->m[i]; a++) {
                    double *g = G + (size_t)(h->nb[i] + a) * nc;
                    for (int k = 0; k < r; k++) {
                        const double u = U[(size_t)a * r + k];
                        for (int e = 0; e < nc; e++) g[e] += u * Zc[i][(size_t)k * nc + e];
                    }
                }
            } else {
                int a = 0;
                for (int c = 0; c < h->nch[i]; c++) {
                    const int j = h->ch0[i] + c;
                    for (int t = 0; t < h->rank[j]; t++, a++)
                        for (int k = 0; k < r; k++) {
                            const double u = U[(size_t)a * r + k];
                            for (int e = 0; e < nc; e++) Zc[j][(size_t)t * nc + e] += u * Zc[i][(size_t)k * nc + e];
                        }
                }
            }
        }
    }
    #pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < nn; i++) {
        if (h->nch[i] != 0) continue;
        for (int f = h->near_ptr[i]; f < h->near_ptr[i + 1]; f++) {
            const int j = h->near_idx[f];
            for (int a = h->nb[i]; a < h->ne[i]; a++) {
                const double *ya = h->Ys + (size_t)a * dim;
                double *g = G + (size_t)a * nc;
                for (int b = h->nb[j]; b < h->ne[j]; b++) {
                    const double *yb = h->Ys + (size_t)b * dim;
                    const double k = kern(ya, yb, dim);
                    g[0] += k;
                    for (int d = 0; d < dim; d++) g[1 + d] += k * yb[d];
                }
            }
        }
    }
    /* s, v in the caller's order (the self term K_ii = 1 removed), then Z and rep */
    double *s = (double *)malloc(sizeof(double) * (size_t)n);
    double *v = (double *)malloc(sizeof(double) * (size_t)n * dim);
    for (int a = 0; a < n; a++) {
        const int i = h->perm[a];
        const double *g = G + (size_t)a * nc;
        s[i] = g[0] - 1.0;
        for (int d = 0; d < dim; d++) v[(size_t)d * n + i] = g[1 + d] - h->Ys[(size_t)a * dim + d];
    }
    double z1 = 0.0, z2 = 0.0, z3 = 0.0;
    for (int i = 0; i < n; i++) {
        double ny = 0.0, yv = 0.0;
        for (int d = 0; d < dim; d++) {
            const double y = Y[(size_t)d * n + i];
            ny += y * y;
            yv += y * v[(size_t)d * n + i];
        }
        z1 += s[i]; z2 += ny * s[i]; z3 += yv;
    }
    const double Z = z1 + 2.0 * z2 - 2.0 * z3;
    if (rep) {
        const double c = 4.0 / Z;
        for (int d = 0; d < dim; d++)
            for (int i = 0; i < n; i++)
                rep[(size_t)d * n + i] = c * (Y[(size_t)d * n + i] * s[i] - v[(size_t)d * n + i]);
    }
    free(s); free(v); free(G);
    for (int i = 1; i < nn; i++) { free(W[i]); free(Zc[i]); }
    free(W); free(Zc);
    return Z;
}

/* Node, leaf and pair counts and the largest / mean skeleton rank of a build. */
void oracle_h2_stats(void *hv, int *n_nodes, int *n_leaves, long long *n_far,
                     long long *n_near, int *rank_max, double *rank_mean)
{
    oh2_t *h = (oh2_t *)hv;
    int nl = 0, rm = 0;
    double rs = 0.0;
    for (int i = 0; i < h->nn; i++) {
        if (h->nch[i] == 0) nl++;
        if (h->rank[i] > rm) rm = h->rank[i];
        rs += h->rank[i];
    }
    if (n_nodes) *n_nodes = h->nn;
    if (n_leaves) *n_leaves = nl;
    if (n_far) *n_far = h->n_far;
    if (n_near) *n_near = h->n_near;
    if (rank_max) *rank_max = rm;
    if (rank_mean) *rank_mean = h->nn > 1 ? rs / (h->nn - 1) : 0.0;
}

/* oracle_run (tsne_oracle.c) with the H2 product on iteration 0 and on rebuild iterations
 * (a fresh build each time, src/tsne_h2pack.c:1765-1776) and the force reused between
 * them, as the reference's H2 path does (:1838-1850); below 2000 points the exact sum
 * every iteration, as the reference.  h2_tol gives the ID tolerance 0.5 h2_tol.  Returns
 * the final KL; *n_builds (if not NULL) receives the builds made. */
double oracle_run_h2(int n, int dim, const int *row_ptr, const int *col_idx, const double *val,
                     double *Y, double *uY, double *gains, int n_iter, double lr,
                     double mom0, double mom1, int mom_switch, double exag0, int exag_iter,
                     int freq, int freq_early, int it_begin, int leaf, double h2_tol,
                     int *n_builds)
{
    const size_t m = (size_t)n * dim;
    double *grad = (double *)malloc(sizeof(double) * m);
    double *rep = (double *)malloc(sizeof(double) * m);
    const int reuse = (n >= 2000);
    int rep_valid = 0, builds = 0;
    double rep_Z = 1.0, kl = 0.0;
    for (int it = it_begin; it < n_iter; it++) {
        const double mom = (it < mom_switch) ? mom0 : mom1;
        const double exag = (it < exag_iter) ? exag0 : 1.0;
        const int rebuilt = (it == it_begin) || oracle_is_rebuild_iter(it, freq, freq_early, exag_iter, reuse);
        memset(grad, 0, sizeof(double) * m);
        oracle_attractive(n, dim, row_ptr, col_idx, val, Y, exag, grad);
        double Z;
        if (reuse && !rebuilt && rep_valid) {
            Z = rep_Z;
        } else if (!reuse) {
            Z = oracle_repulsive_exact(n, dim, Y, rep, NULL, NULL);
        } else {
            void *h = oracle_h2_build(n, dim, Y, leaf, 0.5 * h2_tol);
            Z = oracle_h2_repulsive(h, Y, rep);
            oracle_h2_free(h);
            builds++;
            rep_Z = Z;
            rep_valid = 1;
        }
        for (size_t i = 0; i < m; i++) grad[i] -= rep[i];
        if (it == n_iter - 1) kl = oracle_kl(n, dim, row_ptr, col_idx, val, Y, Z, exag);
        oracle_update(n, dim, Y, uY, gains, grad, mom, lr);
    }
    free(grad); free(rep);
    if (n_builds) *n_builds = builds;
    return kl;
}
