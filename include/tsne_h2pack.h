/*
 * tsne_h2pack.h -- C ABI of the MI355X-native t-SNE gradient library
 * (libtsne_h2pack.so, hand-written gfx950 HIP kernels behind plain C entry points).
 *
 * The thirteen entry points marked [ref] are the ones the reference's ctypes layer
 * binds (tsne_h2pack/_backend.py:345-448 of xinxing02/tsne-h2pack-python 3.3.1); their
 * signatures, struct layouts, ownership and error conventions are those of the
 * reference header include/tsne_h2pack.h, so the reference's Python package can load
 * this library unchanged.  Entry points marked [add] have no counterpart in the
 * reference; they are additive and the thirteen do not depend on them.
 *
 * Conventions (as the reference): every real is float64 at the ABI; X is pt_dim x N and
 * Y is low_dim x N, dimension-major (element d*N+i); failures return NULL or a nonzero
 * code and print a message to stderr.  Device state is private to a workspace: every
 * workspace owns its own HIP stream, so workspaces on one device may run from several
 * host threads at once.
 */
#ifndef TSNE_H2PACK_AMD_H
#define TSNE_H2PACK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR matrix, 40 bytes; reference include/tsne_h2pack.h:17-24.  Created by
 * sparse_matrix_create (malloc) and released by sparse_matrix_destroy (free). */
typedef struct {
    int n_rows;
    int n_cols;
    int nnz;
    int *row_ptr;      /* n_rows + 1 */
    int *col_idx;      /* nnz */
    double *values;    /* nnz */
} sparse_matrix_t;

/* Run parameters, 160 bytes with the reference's field order and offsets
 * (reference include/tsne_h2pack.h:27-85, mirrored by _backend.py:53-95). */
typedef struct {
    int n_iter;                    /*   0 */
    double perplexity;             /*   8 */
    double learning_rate;          /*  16  < 0: max(N/12, 200) */
    double momentum;               /*  24  final momentum */
    double initial_momentum;       /*  32 */
    int momentum_switch_iter;      /*  40 */
    double early_exag;             /*  48 */
    int early_exag_iter;           /*  56 */
    int use_vptree;                /*  60  legacy k-NN selector, unused on this path */
    double h2_tol;                 /*  64  relative tolerance of the H2 approximation */
    int h2_rebuild_freq;           /*  72 */
    int h2pack_print_level;        /*  76 */
    int init_method;               /*  80  0: random N(0, 1e-4); 1: copy init_Y */
    int random_seed;               /*  84 */
    int verbose;                   /*  88 */
    int print_freq;                /*  92 */
    int enable_early_stopping;     /*  96 */
    double convergence_tol;        /* 104 */
    int convergence_patience;      /* 112 */
    int min_iter;                  /* 116 */
    int sklearn_compatible;        /* 120 */
    int enable_profiling;          /* 124 */
    char *profiling_output_file;   /* 128 */
    int num_threads;               /* 136  host threads for host-side work */
    int h2_leaf_size;              /* 140  max points per H2 leaf */
    int h2_rebuild_freq_early;     /* 144 */
    const double *init_Y;          /* 152  low_dim x N, read only inside tsne_init* */
} tsne_params_t;

/* Opaque per-job state (device arrays, H2 structure, stream).  The reference
 * exposes its struct (include/tsne_h2pack.h:102-157) but its Python layer treats it as
 * opaque (_backend.py:123-127). */
typedef struct tsne_workspace tsne_workspace_t;

/* [ref] include/tsne_h2pack.h:167 / src/tsne_h2pack.c:56-106: defaults by value. */
tsne_params_t tsne_create_default_params(void);

/* [ref] include/tsne_h2pack.h:178-184 / src/tsne_h2pack.c:1152: legacy features path.
 * The k-NN search of that path is outside this library's scope: it returns NULL with
 * a message directing callers to compute_P_from_knn_distances + tsne_init_with_P. */
tsne_workspace_t *tsne_init(const double *X, int n_point, int pt_dim, int low_dim,
                            const tsne_params_t *params);

/* [ref] include/tsne_h2pack.h:196-203 / src/tsne_h2pack.c:1356-1547.  Takes ownership of
 * P (released with sparse_matrix_destroy by tsne_destroy) when it returns a workspace;
 * low_dim must be 2 or 3; P must be n_point x n_point.  Uploads P, initialises Y
 * (copy of init_Y or seeded Box-Muller), and for N >= 2000 builds the first H2
 * structure on the device.  NULL on failure; on every NULL return P is left untouched
 * and still belongs to the caller (the reference frees P when its initial H2 build
 * fails, src/tsne_h2pack.c:1518-1523, and keeps it on earlier failures; a caller that
 * frees P after NULL, as the reference's own Python layer should, is safe here). */
tsne_workspace_t *tsne_init_with_P(const double *X, int n_point, int pt_dim, int low_dim,
                                   const tsne_params_t *params, sparse_matrix_t *P);

/* [ref] include/tsne_h2pack.h:210 / src/tsne_h2pack.c:1652-2102: the optimisation loop.
 * Blocking.  0 on success, -1 on failure (non-finite embedding, device error). */
int tsne_run(tsne_workspace_t *workspace);

/* [ref] include/tsne_h2pack.h:217 / src/tsne_h2pack.c:2108-2122: Y_out is low_dim x N in
 * the caller's point order. */
void tsne_get_embedding(tsne_workspace_t *workspace, double *Y_out);

/* [ref] include/tsne_h2pack.h:223 / src/tsne_h2pack.c:2128-2184: frees all, including P. */
void tsne_destroy(tsne_workspace_t *workspace);

/* [ref] include/tsne_h2pack.h:235-242 / src/tsne_h2pack.c:2190-2212 (via tsne_init). */
int tsne_fit_transform(const double *X, int n_point, int pt_dim, int low_dim,
                       const tsne_params_t *params, double *Y_out);

/* [ref] include/tsne_h2pack.h:254-262 / src/tsne_h2pack.c:2214-2237. */
int tsne_fit_transform_with_P(const double *X, int n_point, int pt_dim, int low_dim,
                              const tsne_params_t *params, sparse_matrix_t *P,
                              double *Y_out);

/* [ref] include/tsne_h2pack.h:275 / src/tsne_utils.c:18-39. */
sparse_matrix_t *sparse_matrix_create(int n_rows, int n_cols, int nnz);

/* [ref] include/tsne_h2pack.h:281 / src/tsne_utils.c:41-49. */
void sparse_matrix_destroy(sparse_matrix_t *mat);

/* [ref] include/tsne_h2pack.h:336-344 / src/tsne_utils.c:1022-1105: conditional P_j|i
 * from CSR k-NN L2 distances (not squared) by per-row perplexity bisection, computed on
 * the GPU in FP64 with one wavefront per row.  Returns a malloc'd CSR of the input's
 * shape (free with sparse_matrix_destroy), NULL on failure.  num_threads is accepted
 * for ABI compatibility. */
sparse_matrix_t *compute_P_from_knn_distances(const double *knn_distances,
                                              const int *knn_indices,
                                              const int *knn_indptr, int n_samples,
                                              int nnz, double perplexity, int num_threads);

/* [ref] include/tsne_h2pack.h:430-434 / src/tsne_h2pack.c:1634-1646. */
void tsne_set_init_profiling_times(tsne_workspace_t *workspace, double knn_time,
                                   double p_construction_time);

/* [ref] include/tsne_h2pack.h:405 / src/tsne_h2pack.c:2243-2247: last stored KL. */
double compute_kl_divergence(tsne_workspace_t *workspace);

/* ------------------------------------------------------------------------------------
 * [add] Entry points with no counterpart in the reference.
 * ---------------------------------------------------------------------------------- */

/* Precision of the H2 product.  PLAIN: every stage in FP32 (far-field GEMMs on FP32
 * VALU).  MIXED: near field FP32, far-field basis and coupling products in BF16 with
 * FP32 accumulation on MFMA.  The skeleton selection (interpolative decomposition) is
 * FP64 in both.  Default PLAIN, or the value of the environment variable
 * TSNE_H2_PRECISION ("plain" / "mixed") read at tsne_init_with_P. */
#define TSNE_H2_PLAIN 0
#define TSNE_H2_MIXED 1
int tsne_set_h2_precision(tsne_workspace_t *workspace, int mode);

/* Number of points, embedding dimension and the effective learning rate. */
int tsne_workspace_info(tsne_workspace_t *workspace, int *n_point, int *low_dim,
                        double *learning_rate);

/* Replace the current embedding (low_dim x N, caller's order); momentum and gains are
 * left as they are.  0 on success. */
int tsne_set_embedding(tsne_workspace_t *workspace, const double *Y);

/* Gradient of the current embedding without updating it: the attractive force at
 * exaggeration `exag` plus the repulsive force from a fresh H2 build and product (or
 * the direct sum where the reference uses it).  grad_out is low_dim x N (caller's
 * order), *Z_out the normalisation.  exact != 0 evaluates the repulsive part by the
 * FP64 all-pairs sum on the device instead.  0 on success. */
int tsne_compute_gradient(tsne_workspace_t *workspace, double exag, int exact,
                          double *grad_out, double *Z_out);

/* FP64 exact repulsive force (4/Z) sum_j K_ij (y_i - y_j) on selected rows of the
 * current embedding, with Z from the FP64 all-pairs sum on the device.  out is
 * n_rows x low_dim row-major; *Z_out receives Z.  0 on success. */
int tsne_exact_repulsive_rows(tsne_workspace_t *workspace, const int *rows, int n_rows,
                              double *out, double *Z_out);

/* [add] No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * tsne_exact_repulsive_rows with the same kernels and the same result, bit for bit, but
 * no allocation of any kind.  Its partial results (the FP64 sums of all points, the
 * selected rows, the caller-to-current row map) live in one static 48 MiB device array
 * per GPU, so N * (12 + 8 * low_dim) + n_rows * (16 + 8 * low_dim) bytes (each part
 * rounded up to 256) must fit in 48 MiB -- e.g. N up to about 1.6 million in 2D with
 * 10,000 rows; larger requests fail with a message.  The inputs (rows, a host array)
 * and the output (out, n_rows x low_dim, host) are the caller's.  All calls on one GPU
 * share that array: calls that overlap on one GPU race.  0 on success, -1 otherwise. */
int tsne_exact_repulsive_rows_noalloc(tsne_workspace_t *workspace, const int *rows,
                                      int n_rows, double *out, double *Z_out);

/* Statistics of the last H2 build: levels, nodes, leaves, admissible and near pairs,
 * maximum and mean skeleton rank, proxy points of the finest level, flop counts of one
 * product; of the last mixed-precision product, the nodes whose BF16 coupling was redone
 * in FP32; and the number of builds so far that could not be used (the product then
 * takes the direct sum).  0 on success, -1 if no H2 structure exists. */
typedef struct {
    int levels, n_nodes, n_leaves, leaf_points_max;
    long long n_adm_pairs, n_near_pairs, near_point_pairs;
    int rank_max;
    double rank_mean;
    int proxy_max;
    double far_flops, near_flops;
    int mixed_fallback_nodes;
    int direct_fallbacks;  /* builds so far whose product used the direct sum instead */
    int rank_capped_builds; /* builds so far in which some node's skeleton rank reached
                               the cap (32 in 2D, 48 in 3D) before the ID tolerance; the
                               first one also prints a warning on stderr */
    double id_flops;        /* FP64 work of the last build's IDs: sum over nodes of
                               (3 low_dim + 3) m p + 4 m p r (m candidates, p proxies,
                               r rank; single-child nodes excluded) */
    double id_ms;           /* duration of the last build's ID launches (ms), measured
                               with HIP events while timing mode 1 is on; else 0 */
} tsne_h2_stats_t;
int tsne_h2_stats(tsne_workspace_t *workspace, tsne_h2_stats_t *stats);

/* Phase timings of the last tsne_run, in milliseconds summed over iterations, measured
 * with HIP events on the workspace stream: attractive+update, rebuild, product, other.
 * Also the number of rebuilds.  0 on success. */
int tsne_run_timings(tsne_workspace_t *workspace, double *attr_update_ms,
                     double *rebuild_ms, double *product_ms, double *other_ms,
                     int *n_rebuild);

/* Graph mode for later runs: iterations that reuse the repulsive force (no rebuild, no KL,
 * no timing events) are replayed from HIP graphs of their two kernels (attractive force +
 * update, centring), captured on first use per gather-buffer parity and schedule phase;
 * results are bit-identical to the eager launches.  Default off, or TSNE_GRAPH=1 at
 * tsne_init_with_P; the job server turns it on for its jobs (TSNE_GRAPH=0 keeps it off).
 * 0 on success. */
int tsne_set_graph_mode(tsne_workspace_t *workspace, int on);

/* Iterations replayed from graphs so far (-1 for a NULL workspace). */
int tsne_graph_launches(tsne_workspace_t *workspace);

/* [add] Launch shapes of a job, and their autotuner.  No counterpart in the reference
 * (a CPU library has no launch shapes).  None of these choices changes a result: a job
 * gives the same embedding, bit for bit, under every one of them.
 *   overlap          1: on rebuild iterations the attractive pass runs on a second stream
 *                    beside the rebuild (default: from 2^18 points on)
 *   level_batch      tree levels queued between two host checks (1..8, default 4)
 *   graph            1: force-reusing iterations replayed from HIP graphs (as
 *                    tsne_set_graph_mode; default 0)
 *   one_class_level  internal tree levels of at most this many nodes run their IDs in one
 *                    launch of the widest workgroup shape, larger ones split by size
 *                    class (-1: the default, 512 in 2D, 2048 in 3D)
 *   ms               the tuning run's time (table entries; ignored by tsne_set_tuning) */
typedef struct {
    int overlap;
    int level_batch;
    int graph;
    int one_class_level;
    double ms;
} tsne_tuning_t;

/* Sets a workspace's launch shapes for its later iterations and builds (fields outside
 * their range keep the current choice).  tsne_init_with_P applies the device's tuning
 * table entry for the job's (low_dim, round(log2 N)) when one exists; the job server runs
 * its jobs with the entry's level_batch and one_class_level, graph replay on and overlap
 * off (the table is tuned one job at a time; jobs served together already fill the GPU's
 * queues); environment variables TSNE_OVERLAP / TSNE_GRAPH, when set, take precedence.
 * 0 on success. */
int tsne_set_tuning(tsne_workspace_t *workspace, const tsne_tuning_t *tuning);

/* The tuning table entry for a job of n_point points in low_dim dimensions: 1 and the
 * entry if the device's table holds one for its size bucket, 0 and the built-in defaults
 * otherwise, -1 for invalid arguments. */
int tsne_tuning_lookup(int low_dim, int n_point, tsne_tuning_t *out);

/* Autotuner: times n_iter iterations (the front end's schedule: h2_tol 1e-2 in 2D, 3e-2 in
 * 3D, rebuild every 3, 10 during early exaggeration, leaf 100) of a synthetic job of
 * n_point points (a seeded 10-cluster mixture in 16 dimensions, its exact k = 90 graph and
 * joint P at perplexity 30) under each launch shape, one choice at a time from the
 * defaults, two runs each; checks that every run gives the first run's embedding bit for
 * bit; stores the fastest shapes in the device's table (persisted in the cache directory,
 * as tsne_proxy_cache_info describes, file tune_amd_v1_<arch>_<CUs>cu.txt) and returns
 * them in *best.  n_point >= 2000 (the H2 path), n_iter >= 50.  0 on success. */
int tsne_autotune(int low_dim, int n_point, int n_iter, tsne_tuning_t *best);

/* Path of the device's tuning table file; -1 (and "") when the cache is off or no device
 * is present. */
int tsne_tuning_path(char *path, int path_len);

/* [add] Diagnostics, no counterpart in the reference: average time (ms, HIP events on the
 * workspace stream, over `reps` launches after one warm-up) of the attractive pass's memory
 * pattern over the workspace's current P and embedding.  mode 0: P streamed only; 1: P and
 * the neighbour gathers (sum of p*y_j); 2: the full attractive sums without the update.
 * The fused attractive/update kernel's time against modes 0 and 1 says how close it runs
 * to its own access pattern's ceiling.  Leaves the embedding unchanged. */
int tsne_attr_probe(tsne_workspace_t *workspace, int mode, int reps, double *avg_ms);

/* State of the workspace's run so far: H2 rebuilds inside the loop since iteration 0
 * (the initial build of tsne_init_with_P not counted), point renumberings, the Z of the
 * last repulsive product (the one the KL of the last iteration uses) and the last KL
 * evaluated (at the last iteration, and every 50th when verbose).  Any pointer may be
 * NULL.  0 on success. */
int tsne_run_state(tsne_workspace_t *workspace, int *n_rebuild, int *n_renumber, double *Z,
                   double *kl);

/* Run only iterations [it_begin, it_end) of the schedule of an n_iter-iteration run
 * (the schedule follows the absolute iteration number).  tsne_run equals
 * tsne_run_range(ws, 0, n_iter).  Lets a caller time a window of a run. */
int tsne_run_range(tsne_workspace_t *workspace, int it_begin, int it_end);

/* Block until every operation queued on the workspace stream is done. */
int tsne_synchronize(tsne_workspace_t *workspace);

/* Timing of later runs.  mode 1: per-phase HIP events, synchronising every iteration
 * (read with tsne_run_timings; also on when params.enable_profiling is set).  mode 2:
 * HIP events on the workspace stream around every attractive+update kernel launch,
 * read once after the run (read with tsne_attr_kernel_time).  0 turns both off.
 * 0 on success. */
int tsne_enable_timing(tsne_workspace_t *workspace, int mode);

/* Summed duration (ms) and number of the attractive+update kernel launches timed in
 * the last run under timing mode 2.  0 on success. */
int tsne_attr_kernel_time(tsne_workspace_t *workspace, double *total_ms, int *launches);

/* The part of tsne_attr_kernel_time spent in launches of the tiled attractive kernel
 * (after the first renumbering) and their number.  0 on success. */
int tsne_attr_kernel_time_tiled(tsne_workspace_t *workspace, double *total_ms, int *launches);

/* Row tiles of the attractive pass after the last renumbering: rows per tile, tile count,
 * largest distinct-neighbour count among the tiles staged in LDS, total distinct
 * neighbours over all tiles, and tiles whose neighbours are read from HBM instead
 * (ntiles 0: no renumbering yet, the per-row kernel runs).  0 on success. */
int tsne_attr_tile_stats(tsne_workspace_t *workspace, int *tile_rows, int *ntiles, int *umax,
                         long long *ucount, int *n_global);

/* [add] Joint P from k-NN distances on the device: the conditional P of
 * compute_P_from_knn_distances, then P = Pc + Pc^T and P /= sum(P) -- the symmetrisation
 * the reference does on the host with scipy (tsne_h2pack/_backend.py:695-718), bit for
 * bit: entries summing to exactly 0 are dropped, the sum is scipy's (rows summed in
 * order, then numpy's blocked pairwise sum).  Rows must have increasing columns in
 * [0, n) and no self entries (the reference's neighbour indexing).  Returns a matrix from
 * sparse_matrix_create (release with sparse_matrix_destroy) or NULL. */
sparse_matrix_t *tsne_joint_P_from_knn_distances(const double *knn_distances,
                                                 const int *knn_indices, const int *knn_indptr,
                                                 int n_samples, int nnz, double perplexity);

/* [add] numpy's float64 sum of n contiguous values (blocks of 8192 summed pairwise, in
 * order), the normalisation the joint P uses. */
double tsne_numpy_sum(const double *values, long long n);

/* [add] Exact Euclidean k-NN graph on the device.  Replaces the reference's exact search
 * (tsne_h2pack/_knn_methods.py:71-119, ExactKNN: sklearn NearestNeighbors(algorithm=
 * 'brute').kneighbors_graph(mode='distance'), data cast to float32).  X is n_samples x
 * n_features float64, row-major.  Row i of indices / distances (n_samples x k) holds the k
 * nearest other points in increasing distance (ties by index) and their float32 L2
 * distances; the point itself is removed as sklearn removes it (or the nearest entry when
 * more than k duplicates precede it).  Neighbour sets are exact in FP64: a BF16-split MFMA
 * screening keeps candidates within a rigorous error bound, FP64 ranks them, and rows the
 * bound cannot settle are re-ranked in FP64.  1 <= k < n_samples, k <= 900; X finite.
 * stats may be NULL.  0 on success, -1 otherwise (message on stderr). */
typedef struct {
    int survivors;       /* candidates per point kept from the screening */
    int uncertain_rows;  /* rows re-ranked by the FP64 fallback */
    int overflow_rows;   /* of those, rows ranked by a full sort of their distances */
    double eps;          /* screening error bound relative to |x|^2 + |y|^2 (centred) */
    double device_ms;    /* device time, first preparation kernel to the last kernel */
    double screen_ms;    /* of which the MFMA screening launches */
    long long screen_flops;  /* BF16 MFMA flops of the screening */
} tsne_knn_stats_t;
int tsne_knn_exact(const double *X, int n_samples, int n_features, int k, int *indices,
                   float *distances, tsne_knn_stats_t *stats);

/* [add] Joint P from the points, entirely on the device: the exact k-NN graph of
 * tsne_knn_exact (float32 distances, as the reference's exact provider returns them), the
 * reference's neighbour indexing (columns sorted per row, distances to float64;
 * tsne_h2pack/_backend.py:647-674), then tsne_joint_P_from_knn_distances.  The result is
 * the joint P those three steps give on the host, bit for bit.  Returns a matrix from
 * sparse_matrix_create (release with sparse_matrix_destroy) or NULL.  stats may be NULL. */
sparse_matrix_t *tsne_joint_P_from_points(const double *X, int n_samples, int n_features, int k,
                                          double perplexity, tsne_knn_stats_t *stats);

/* [add] Embedding quality on a common footing, on the device: the reference's score()
 * (benchmarks/quality.py:56-81).  Y: n x dim float64 row-major embedding; hd_idx: n x hd_k
 * high-dimensional neighbours per point in increasing distance (the shared k-NN graph, self
 * removed); queries / anchors: point indices (the reference draws them with
 * numpy default_rng(0).choice); P: the common joint P (CSR, int32 row pointers).
 *   knn_preservation = mean over queries of |hd[i, :k] & ld(i)| / k, ld(i) = the k+1
 *     nearest embedding points of y_i (exact, FP64; ties by index) without the first;
 *   Z = n * mean over anchors a of sum_{j != a} 1 / (1 + |y_a - y_j|^2);
 *   kl = sum p log p + sum p log1p(|y_i - y_j|^2) + log Z over the nonzeros of P;
 *   span = the largest coordinate range.
 * 1 <= k <= 15, k <= hd_k.  0 on success, -1 otherwise (message on stderr). */
typedef struct {
    double kl;
    double knn_preservation;
    double Z;
    double span;
} tsne_quality_t;
int tsne_quality_score(const double *Y, int n, int dim, const int *hd_idx, int hd_k, int k,
                       const int *queries, int n_queries, const int *P_row_ptr,
                       const int *P_col_idx, const double *P_values, const int *anchors,
                       int n_anchors, tsne_quality_t *out);

/* [add] Persistent proxy-point cache (host only, no device needed).  The reference keeps
 * its proxy file across runs (src/tsne_h2pack.c:1093-1146; TSNE_H2PACK_CACHE_DIR, else
 * $XDG_CACHE_HOME/tsne_h2pack, else $HOME/.cache/tsne_h2pack; TSNE_H2PACK_NO_PROXY_CACHE=1
 * turns it off); this library does the same with its own versioned binary format (magic
 * "TSNEPRXY", version, table parameters, counts, points, FNV-1a checksum).  Loads or
 * generates (and stores) the proxy table an H2 build with (low_dim, h2_tol) uses, and
 * reports the cache file path ("" when off), a digest of the table and whether it was read
 * from disk.  0 on success, -1 for invalid arguments. */
int tsne_proxy_cache_info(int low_dim, double h2_tol, char *path, int path_len,
                          unsigned long long *digest, int *from_disk);

/* ------------------------------------------------------------------------------------
 * [add] Job server: continuous batching of independent jobs on one GPU.  No counterpart
 * in the reference (one job per process there).  Jobs are admitted in arrival order
 * into `slots` concurrent slots while their estimated device footprint fits the HBM
 * budget; each runs on its own HIP stream with its state resident on the device, and a
 * finished job's slot and memory go to the next queued job at once.  A job's result is
 * the same as that of tsne_init_with_P + tsne_run + tsne_get_embedding on the same
 * inputs.
 * ---------------------------------------------------------------------------------- */
#define TSNE_JOB_QUEUED 0
#define TSNE_JOB_RUNNING 1
#define TSNE_JOB_DONE 2
#define TSNE_JOB_FAILED (-1)

typedef struct tsne_server tsne_server_t;

typedef struct {
    int job_id;
    int status;                /* TSNE_JOB_* */
    int n_point, low_dim;
    double kl;                 /* final KL divergence (done jobs) */
    double submit_s, start_s, end_s;  /* seconds since the server was created */
    long long footprint_bytes; /* device memory reserved for the job at admission */
} tsne_job_info_t;

/* Estimated peak device bytes of one job (the admission charge). */
long long tsne_job_footprint(int n_point, int low_dim, long long nnz);

/* slots: concurrent jobs (1..64); hbm_budget_bytes <= 0: 90% of the free device memory.
 * NULL without a device. */
tsne_server_t *tsne_server_create(int slots, long long hbm_budget_bytes);

/* Queue a job: P is the joint P (n_point x n_point, from sparse_matrix_create), taken by
 * the server on success; init_Y (low_dim x n_point, may be NULL for the seeded random
 * init) is copied.  Returns the job id (>= 0), or -1 (P not taken). */
int tsne_server_submit(tsne_server_t *server, int n_point, int low_dim,
                       const tsne_params_t *params, sparse_matrix_t *P, const double *init_Y);

/* H2 product precision (0 plain, 1 mixed, as tsne_set_h2_precision) of the jobs submitted
 * after this call; until it is called, each job keeps the workspace default
 * (TSNE_H2_PRECISION).  0 on success. */
int tsne_server_set_h2_precision(tsne_server_t *server, int mode);

/* Status of a job (TSNE_JOB_*), -2 for an unknown id; fills *info if not NULL. */
int tsne_server_poll(tsne_server_t *server, int job_id, tsne_job_info_t *info);

/* Block until the job ends; copy its embedding (low_dim x n_point, caller order) to Y_out
 * if not NULL, fill *info, and release the job.  0 if it succeeded, -1 otherwise.  One
 * waiter per job: while a wait on job_id is in progress (or after it returned), another
 * wait on the same id returns -1 at once. */
int tsne_server_wait(tsne_server_t *server, int job_id, double *Y_out, tsne_job_info_t *info);

/* Queued and running job counts and the device bytes charged to running jobs. */
int tsne_server_status(tsne_server_t *server, int *queued, int *running,
                       long long *hbm_in_use);

/* Drops queued jobs (they end as failed), waits for running ones, wakes every
 * tsne_server_wait still blocked in the server and waits until each has returned, then
 * frees everything.  No call may use the server once destroy has begun. */
void tsne_server_destroy(tsne_server_t *server);

#ifdef __cplusplus
}
#endif

#endif /* TSNE_H2PACK_AMD_H */
