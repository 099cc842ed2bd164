// kernels.h -- launchers of the per-iteration kernels (kernels_iter.hip) and the
// device scalar slots they share.
#pragma once
#include <hip/hip_runtime.h>

namespace tsne {

// Slots of the per-workspace device scalar array `scal` (double).
// The attractive pass reads P's columns and fp32 values in aligned 4- and 8-element
// chunks, so those two arrays carry this many elements of padding past nnz.
constexpr int kCsrPad = 8;

enum ScalarSlot {
    SLOT_Z = 0,        // normalisation Z of the current repulsive force
    SLOT_MEAN = 1,     // 1..3: pending centring mean of Y64 (removed at next read)
    SLOT_YSHIFT = 4,   // 4..6: the mean in force when the repulsive product was formed
    SLOT_BAD = 7,      // non-finite embedding seen
    SLOT_ZBAD = 8,     // non-positive / non-finite Z seen (replaced by 1)
    SLOT_KL = 9,       // last KL divergence
    SLOT_ZEXACT = 10,  // Z of the FP64 check path
    SLOT_ZSKIP = 11,   // 1 while the last product's Z was non-positive / non-finite: the
                       // updates then leave the repulsive force out, as the reference
                       // leaves it out of the gradient (src/tsne_h2pack.c:604-611)
    SLOT_COUNT = 16
};

// Scale of the unscaled repulsive sum F in the gradient: 4 / Z, or 0 while SLOT_ZSKIP is set.
__device__ __forceinline__ double rep_scale_of(const double *scal)
{
    return scal[SLOT_ZSKIP] != 0.0 ? 0.0 : 4.0 / scal[SLOT_Z];
}

int launch_attr_update(int D, int n, const int *row_ptr, const int *col_idx, const float *pval,
                       const void *Yg, const float *F, double *Y64, double *uY, double *gains,
                       void *Yg_next, const double *scal, double exag, double mom, double lr,
                       double *block_sums, int nblocks, hipStream_t st,
                       hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
struct AttrTiles;
// The same pass over row tiles with LDS-staged neighbours (tiles.h), nt (256 or 1024)
// threads per tile; nblocks is the capacity of block_sums in records.
int launch_attr_tile(int D, int n, const AttrTiles &tiles, int nt, const int *row_ptr,
                     const int *col_idx, const float *pval, const void *Yg, const float *F,
                     double *Y64, double *uY, double *gains, void *Yg_next, const double *scal,
                     double exag, double mom, double lr, double *block_sums, int nblocks,
                     hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
// The attractive pass split in two, for iterations whose repulsive force is rebuilt
// meanwhile on another stream: the attractive sums into A (float [n][D]), then the
// update from A and the repulsive sum F (nblocks as for launch_attr_update).  grid: the
// attractive launch's workgroups (0: nblocks; fewer leave CUs to the rebuild's launches,
// with the same sums).
int launch_attr_only(int D, int n, const int *row_ptr, const int *col_idx, const float *pval,
                     const void *Yg, float *A, int nblocks, int grid, hipStream_t st);
int launch_update(int D, int n, const float *A, const float *F, double *Y64, double *uY,
                  double *gains, void *Yg_next, const double *scal, double exag, double mom,
                  double lr, double *block_sums, int nblocks, hipStream_t st);
// Workgroups of the first stage of the centring-mean reduction: block-sum buffers hold
// kMeanParts records (3 doubles each) beyond the blocks of the attractive launch.
constexpr int kMeanParts = 256;
// Diagnostics: the attractive pass's memory pattern; mode 0 P only, 1 P and the
// neighbour gathers, 2 the full attractive sums (launch_attr_only).  A: float [n][D].
int launch_attr_probe(int D, int n, int mode, const int *row_ptr, const int *col_idx,
                      const float *pval, const void *Yg, float *A, int nblocks, hipStream_t st);
int launch_repulsive_direct_f32(int D, int n, const void *Yg, float *s, float *F,
                                hipStream_t st);
int launch_z(int D, int n, const void *Yg, const float *s, const float *F, double *scal,
             double *partial, int nblocks, hipStream_t st);
int launch_shift_snapshot(double *scal, hipStream_t st);
int launch_f64_to_f32(const double *src, float *dst, size_t n, hipStream_t st);
int launch_fill_f64(double *dst, double v, size_t n, hipStream_t st);
int launch_kl(int D, int n, const int *row_ptr, const int *col_idx, const double *pval64,
              const double *Y64, const double *scal, double exag, double *row_kl,
              double *partial, int nblocks, double *dst, hipStream_t st);
int launch_repulsive_direct_f64(int D, int n, const double *Y64, const int *rows, int n_rows,
                                double *s, double *F, hipStream_t st);
int launch_z_f64(int D, int n, const double *Y64, const double *s, const double *F,
                 const double *scal, double *partial, int nblocks, double *dst,
                 hipStream_t st);
int launch_perplexity(int n, const double *dist, const int *indptr, double perplexity,
                      double *p_out, hipStream_t st);

}  // namespace tsne
