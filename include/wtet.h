/*
 * wtet_amd -- C ABI of the MI355X Delaunay triangulation path.
 *
 * Plain pointers and sizes only.  Every entry point returns a status code
 * (never aborts the process); wtet_last_error() describes the last failure
 * of a context.  Each declaration names the reference interface it
 * replaces (paths relative to the reference repository).
 */
#ifndef WTET_AMD_H
#define WTET_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WTET_INVALID 0xFFFFFFFFu /* src/types.rs:177 INVALID */

typedef enum wtet_status {
  WTET_OK = 0,
  WTET_ERR_INVALID_ARGUMENT = 1, /* bad shape / null pointer / bad option    */
  WTET_ERR_TOO_FEW_POINTS = 2,   /* n < 4 (src/python.rs:182-187, lib.rs:167) */
  WTET_ERR_DEVICE = 3,           /* HIP runtime failure or no device          */
  WTET_ERR_OUT_OF_MEMORY = 4,    /* device or host allocation failed          */
  WTET_ERR_NOT_CONVERGED = 5,    /* could not complete the triangulation      */
  WTET_ERR_INTERNAL = 6          /* invariant violated inside the library     */
} wtet_status;

/* GDelConfig, src/types.rs:32-61 (defaults :49-61); the three trailing
 * fields are additions of this library. */
typedef struct wtet_config {
  int32_t insertion_rule;         /* 0 circumcenter, 1 centroid (types.rs:17-28) */
  int32_t enable_flipping;        /* 1 */
  int32_t enable_sorting;         /* 0: Morton order requested                 */
  int32_t enable_hilbert_sorting; /* 0: Hilbert + stratified order requested    */
  int32_t enable_splaying;        /* 1: repair any face the GPU leaves non-Delaunay */
  uint32_t max_insert_iterations; /* 100 */
  uint32_t max_flip_iterations;   /* 10: flip iterations per insertion round   */
  int32_t predicate_mode;         /* 0: f32 filter -> f64 filter -> exact; 1: exact only */
  int32_t collect_stats;          /* 1: count predicate stages (small cost)    */
  int32_t presize_mode;           /* tet-slot presizing: 0 fixed, the reference's
                                     max(8(N+5)+8, 12N, 64) (src/gpu/mod.rs:101-108);
                                     1 learned (model set with wtet_set_presize_model);
                                     2 none (grow-and-retry from 2N) */
} wtet_config;

/* DelaunayResult, src/types.rs:69-79, owned by the caller after the call;
 * release with wtet_result_free. */
typedef struct wtet_result {
  uint32_t num_tets;
  uint32_t *tets;      /* num_tets*4 vertex ids: input ids, super n..n+3, infinity n+4 */
  uint32_t *adjacency; /* num_tets*4: (tet << 5) | face of the neighbour across face f,
                          face f opposite vertex f; WTET_INVALID when none */
  uint32_t num_failed;
  uint32_t *failed_verts; /* input ids of vertices opposite faces the GPU phase left
                             non-Delaunay (before the host repair), sorted */
} wtet_result;

typedef struct wtet_stats {
  uint64_t pred[8]; /* orient3d: f32, f64, exact nonzero, exact zero (SoS);
                       insphere: f32, f64 (filter or double-double adaptive
                       stage), exact nonzero, exact zero (SoS) */
  uint64_t sum_uninserted; /* sum over rounds of uninserted points            */
  uint64_t sum_active;     /* sum over flip iterations of tets checked         */
  uint64_t sum_flips;      /* flips executed                                  */
  uint64_t num_points;
  uint64_t num_tets_out;   /* tets in the result (incl. super/infinity tets)   */
  uint64_t tet_slots;      /* tet slots used on the device                     */
  uint32_t rounds;
  uint32_t flip_iterations;
  uint32_t violations_gpu; /* non-Delaunay faces left by the GPU phase          */
  uint32_t repair_regions; /* regions re-triangulated on the host             */
  uint32_t walk_failures;
  uint32_t grow_events;    /* device buffer growths                           */
  uint32_t duplicates;     /* exact duplicate points not inserted (all but the
                              highest id of each group)                       */
  double ms_device;        /* normalized points in HBM -> tets in HBM          */
  double ms_repair;        /* host repair                                      */
  double ms_total;         /* whole call                                       */
  uint64_t peak_active;    /* longest active (flip-check) list                 */
  uint64_t presize_slots;  /* tet slots requested before the loop              */
  uint64_t tet_capacity;   /* tet slots allocated when the loop started        */
  float presize_pred[3];   /* learned presizing: predicted log(slots / n),
                              log(peak active / n), log(1 + stuck faces)        */
  uint64_t repair_rows_staged;   /* rows copied to the host for the repair     */
  uint32_t repair_full_downloads; /* repairs that needed every row            */
  double ms_repair_staging;       /* of ms_repair: rows and points to the host */
  double ms_repair_compute;       /* of ms_repair: the host re-triangulation   */
  uint32_t repair_fetches;        /* rows fetched on demand during the repair */
  uint32_t repair_retries;        /* device phases rerun with a wider insertion
                                     ring because a repair region outgrew its
                                     cap (the stats above are those of the
                                     last run)                                */
  uint32_t repair_max_region;     /* tets in the largest repair region        */
  uint32_t points_left;           /* points not inserted when max_insert_iterations
                                     rounds had run (src/phase1.rs:26); the host
                                     phase inserts them when enable_splaying is set
                                     and failed_verts lists them                */
  uint64_t flips_23;              /* executed 2-3 flips (of sum_flips)         */
  uint64_t flips_32;              /* executed 3-2 flips (of sum_flips)         */
  uint64_t small_checked;         /* of sum_active: checked in the one-workgroup
                                     iterations of short lists              */
  uint64_t small_flips;           /* of sum_flips: executed in those iterations */
} wtet_stats;

typedef struct wtet_ctx wtet_ctx;

/* Device singleton of src/python.rs:13-47 becomes an explicit context. */
wtet_status wtet_create(int device, wtet_ctx **out);
void wtet_destroy(wtet_ctx *ctx);
const char *wtet_last_error(const wtet_ctx *ctx);

void wtet_config_default(wtet_config *cfg);

/* delaunay_3d(device, queue, points, config), src/lib.rs:160-310;
 * xyz: n*3 row-major f32 in host memory. */
wtet_status wtet_delaunay(wtet_ctx *ctx, const float *xyz, uint32_t n,
                          const wtet_config *cfg, wtet_result *out);

/* Same with the points already resident in device memory (d_xyz: n*3 f32). */
wtet_status wtet_delaunay_device(wtet_ctx *ctx, const float *d_xyz, uint32_t n,
                                 const wtet_config *cfg, wtet_result *out);

void wtet_result_free(wtet_result *res);

/* Result left in device memory owned by the context (valid until the next
 * call on the same context): same layout as wtet_result. */
typedef struct wtet_device_result {
  uint32_t num_tets;
  uint32_t *d_tets;
  uint32_t *d_adjacency;
  uint32_t num_failed;
} wtet_device_result;

/* Points and result both resident in device memory: no host copy of the
 * mesh unless the host repair runs (it then downloads the mesh, repairs it
 * and uploads the result).  No counterpart in the reference, whose
 * delaunay_3d always reads back (src/lib.rs:181, src/gpu/mod.rs:207-332). */
wtet_status wtet_delaunay_resident(wtet_ctx *ctx, const float *d_xyz, uint32_t n,
                                   const wtet_config *cfg, wtet_device_result *out);

/* Many independent clouds in one pass: cloud c is the points
 * [offsets[c], offsets[c+1]) of xyz (num_clouds+1 offsets, offsets[0] = 0,
 * at least 4 points per cloud), normalized on its own, and its result is the
 * triangulation wtet_delaunay gives for those points alone (cloud-local
 * vertex ids: input id within the cloud, super n_c..n_c+3, infinity n_c+4).
 * Every kernel launch covers all clouds.  No counterpart in the reference,
 * which triangulates one cloud per delaunay_3d call (src/lib.rs:160-310);
 * the per-cloud semantics are those of that call. */
typedef struct wtet_batch_result {
  uint32_t num_clouds;
  uint64_t num_tets;        /* rows over all clouds */
  uint32_t *tets;           /* num_tets*4 */
  uint32_t *adjacency;      /* num_tets*4: (row within the cloud << 5) | face */
  uint64_t *tet_offsets;    /* num_clouds+1: cloud c owns rows [tet_offsets[c], tet_offsets[c+1]) */
  uint64_t num_failed;
  uint32_t *failed_verts;   /* per cloud, as wtet_result.failed_verts */
  uint64_t *failed_offsets; /* num_clouds+1 */
} wtet_batch_result;

wtet_status wtet_delaunay_batch(wtet_ctx *ctx, const float *xyz, const uint64_t *offsets,
                                uint32_t num_clouds, const wtet_config *cfg,
                                wtet_batch_result *out);
/* Same with the points already in device memory (offsets in host memory). */
wtet_status wtet_delaunay_batch_device(wtet_ctx *ctx, const float *d_xyz, const uint64_t *offsets,
                                       uint32_t num_clouds, const wtet_config *cfg,
                                       wtet_batch_result *out);
void wtet_batch_result_free(wtet_batch_result *res);

/* Points and result both in device memory.  Cloud c owns the device rows
 * [tet_start[c], tet_start[c] + tet_count[c]); rows of different clouds do
 * not overlap but need not be in cloud order (a cloud the host repair
 * rewrote is placed after the others).  Device arrays and the two host
 * arrays are owned by the context and valid until its next call. */
typedef struct wtet_batch_device_result {
  uint32_t num_clouds;
  uint64_t rows;              /* rows used in the device arrays */
  uint32_t *d_tets;
  uint32_t *d_adjacency;
  const uint64_t *tet_start;  /* host, num_clouds */
  const uint32_t *tet_count;  /* host, num_clouds */
  uint64_t num_failed;
} wtet_batch_device_result;

wtet_status wtet_delaunay_batch_resident(wtet_ctx *ctx, const float *d_xyz, const uint64_t *offsets,
                                         uint32_t num_clouds, const wtet_config *cfg,
                                         wtet_batch_device_result *out);

/* Counters of the last wtet_delaunay* call on this context. */
wtet_status wtet_get_stats(const wtet_ctx *ctx, wtet_stats *out);

/* check_quality(points, result), src/python.rs:282-339 (check_delaunay_quality,
 * src/lib.rs:25-78): insphere violations among tet pairs whose vertices are
 * all real, and the number of such interior tets.  Host code. */
wtet_status wtet_check_quality(const float *xyz, uint32_t n, const uint32_t *tets,
                               const uint32_t *adjacency, uint32_t num_tets,
                               uint64_t *violations, uint64_t *interior_tets);

/* encode_adjacency / decode_adjacency, src/python.rs:256-272 */
uint32_t wtet_encode_adjacency(uint32_t tet, uint32_t face);
void wtet_decode_adjacency(uint32_t packed, uint32_t *tet, uint32_t *face);

/* __version__ and gpu_info(), src/python.rs:355-371, 397-399 */
const char *wtet_version(void);
wtet_status wtet_device_info(wtet_ctx *ctx, char *buf, uint32_t len);

/* Device memory for callers without their own HIP allocator (one process
 * should hold one HIP runtime): copy host bytes to a new device allocation,
 * copy device bytes back, release.  Synchronous. */
wtet_status wtet_device_upload(wtet_ctx *ctx, const void *host, size_t bytes, void **d_out);
wtet_status wtet_device_download(wtet_ctx *ctx, const void *d_src, void *host_dst, size_t bytes);
wtet_status wtet_device_free(wtet_ctx *ctx, void *d_ptr);
/* Device-to-device copy of `bytes` (a multiple of 16, 16-byte aligned
 * pointers) by a plain streaming kernel on the context's stream;
 * synchronous.  Also the known-size kernel that HBM byte counters are
 * calibrated against. */
wtet_status wtet_device_copy(wtet_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);

/* The library's stable radix sort (the point order of enable_sorting /
 * enable_hilbert_sorting, which the reference sorts on the host,
 * src/gpu/mod.rs:422-520): n (key, value) pairs in device memory, keys of
 * key_bits 32 or 64 bits sorted on their low end_bit bits, equal keys in
 * input order; inputs are read only and must not overlap the outputs.
 * Synchronous. */
wtet_status wtet_device_sort_pairs(wtet_ctx *ctx, const void *d_keys, const uint32_t *d_vals,
                                   void *d_keys_out, uint32_t *d_vals_out, uint32_t n,
                                   uint32_t key_bits, int end_bit);

/* Points inserted in one round must lie in tets at least `ring`+1
 * face-steps apart (ring 0: one point per tet, as the reference's vote,
 * src/shaders/vote.wgsl:235-344; default 1).  A wider ring leaves fewer
 * faces that concurrent flipping cannot fix, at the cost of more rounds.
 * The triangulation does not depend on it.  No counterpart in the
 * reference. */
wtet_status wtet_set_insertion_ring(wtet_ctx *ctx, int ring);
/* The first `rounds` insertion rounds use ring 0 whatever the ring: while
 * the mesh is small a round costs about the same however few points it
 * inserts, and ring 0 inserts several times as many.  0xFFFFFFFF (the
 * default): round(log4(largest cloud)) - 4, clamped to 0..6.  The
 * triangulation does not depend on it.  No counterpart in the reference. */
wtet_status wtet_set_insertion_warmup(wtet_ctx *ctx, uint32_t rounds);
/* A resident result's host repair works on the rows within `rings`
 * face-steps (1..16, default 7; three more, at most 16, when the stuck
 * faces number over one per 2,000 points) of each stuck face, copied to
 * the host; rows it reaches beyond them are fetched on demand (at most 32
 * times), past that every row is copied.  No counterpart in the
 * reference. */
wtet_status wtet_set_repair_rings(wtet_ctx *ctx, int rings);
/* A single-cloud call whose host repair grows a region past `tets` tets
 * (default 16,384; 0: no limit) stops that repair and reruns the device
 * phase with the insertion ring one wider (at most 2), whose repair then
 * has no limit.  The triangulation does not depend on it; the stats count
 * the reruns (repair_retries).  No counterpart in the reference. */
wtet_status wtet_set_repair_region_cap(wtet_ctx *ctx, uint32_t tets);

/* The audit of a call records at most `records` (>= 64, default 2^20)
 * violated faces; when it finds more it runs again into a buffer as large
 * as the count, so the host repair starts from every violated face.  No
 * counterpart in the reference, whose gather clamps failed_verts to n
 * (src/gpu/mod.rs:306-316). */
wtet_status wtet_set_audit_records(wtet_ctx *ctx, uint32_t records);

/* The flip iterations of a round run on the device without host reads
 * (counts, iteration tag and list parity live in device counters).  Each
 * insertion round after its pick -- splits, points into the children, the
 * round's flip iterations, relocation, and the next round's vote and pick
 * -- is captured in a HIP graph keyed by its grid sizes (powers of two over
 * the host's bounds; every kernel reads its counts from device counters)
 * and replayed for every round and call that fits it, with one host read
 * per round (on by default; off: the same kernels launched directly).  A
 * graph is captured the second time its key comes up: the first time its
 * kernels are launched directly.  The list kernels are persistent grid-stride kernels
 * whose grid is sized per round from the host's bound on the first list,
 * for `items` entries per thread (default 2; powers of two from 16 to 4096
 * workgroups, one graph per size).  No counterpart in the reference, which
 * reads counters back every iteration (src/phase1.rs:16-1431). */
wtet_status wtet_set_flip_graphs(wtet_ctx *ctx, int on);
wtet_status wtet_set_flip_items(wtet_ctx *ctx, uint32_t items);
wtet_status wtet_get_flip_items(const wtet_ctx *ctx, uint32_t *items, uint32_t *graph_captures);
/* Autotuner: triangulates the given clouds (device points) with 1, 2, 4
 * and 8 entries per thread, each setting twice untimed (the second call
 * captures its graphs) and three times timed; keeps the setting with the
 * lowest median device time for this context and reports that median. */
wtet_status wtet_autotune(wtet_ctx *ctx, const float *d_xyz, const uint64_t *offsets,
                          uint32_t num_clouds, const wtet_config *cfg, uint32_t *best_items,
                          double *best_ms);

/* Learned presizing (no counterpart in the reference; the reference sizes
 * its buffers with a fixed formula, src/gpu/buffers.rs:112-579).
 * Features of a cloud: the 512-bin occupancy histogram of the top 9 bits of
 * its 30-bit Morton code (src/morton.rs:22-35), mean and covariance of its
 * box-normalized coordinates, log(n).  The MLP (522 -> 128 -> 128 -> 3,
 * ReLU, FP32) predicts log(tet slots / n), log(peak active list / n) and
 * log(1 + stuck faces); its parameters are one float array of
 * WTET_PRESIZE_PARAMS values: feature mean[522], feature std[522], target
 * mean[3], target std[3], W1[522*128], b1[128], W2[128*128], b2[128],
 * W3[128*3], b3[3], margin[3] (weights input-major). */
#define WTET_PRESIZE_FEATURES 522
#define WTET_PRESIZE_OUTPUTS 3
#define WTET_PRESIZE_PARAMS 84896

/* features of each cloud (d_xyz in device memory, offsets as for
 * wtet_delaunay_batch) into feat (host, num_clouds * 522) */
wtet_status wtet_presize_features(wtet_ctx *ctx, const float *d_xyz, const uint64_t *offsets,
                                  uint32_t num_clouds, float *feat);
/* Same features, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  No counterpart in the
 * reference.  It allocates nothing: the offsets, per-cloud bounding boxes,
 * histogram accumulators and feature staging live in one static device
 * array of WTET_NOALLOC_SCRATCH_BYTES per GPU (resolved by wtet_create,
 * held until the process ends), so at most WTET_NOALLOC_MAX_CLOUDS clouds
 * per call (more: WTET_ERR_INVALID_ARGUMENT).  Every call on one GPU shares
 * that array: calls that overlap on one GPU (two contexts, two threads)
 * race.  Synchronous. */
#define WTET_NOALLOC_MAX_CLOUDS 2048
#define WTET_NOALLOC_SCRATCH_BYTES 8675840
wtet_status wtet_presize_features_noalloc(wtet_ctx *ctx, const float *d_xyz, const uint64_t *offsets,
                                          uint32_t num_clouds, float *feat);
/* GPU inference from host arrays: pred (num_clouds * 3) in log space,
 * margin included */
wtet_status wtet_presize_predict(wtet_ctx *ctx, const float *params, const float *feat,
                                 uint32_t num_clouds, float *pred);
/* model used by presize_mode 1 (null clears it) */
wtet_status wtet_set_presize_model(wtet_ctx *ctx, const float *params);
/* MSE loss and its gradient (trainable range W1..b3 of the parameter
 * array; the rest is zero) over rows of raw features and raw targets,
 * standardized with the parameters' mean/std, computed by the training
 * kernels (rows <= 4096) */
wtet_status wtet_mlp_loss_grad(wtet_ctx *ctx, const float *params, const float *feat,
                               const float *targets, uint32_t rows, float *loss, float *grad);
/* Adam training on the GPU: params (in: initial weights and the
 * standardization; out: trained), minibatches of `batch` rows drawn by a
 * seeded shuffle each epoch; loss_curve (may be null) gets each epoch's
 * mean minibatch loss */
wtet_status wtet_mlp_train(wtet_ctx *ctx, float *params, const float *feat, const float *targets,
                           uint32_t rows, uint32_t epochs, uint32_t batch, float lr,
                           uint64_t seed, float *loss_curve);

/* Per-kernel-group timing with HIP events on the context's stream (the
 * reference's GpuProfiler/CpuProfiler scopes, src/profiler.rs:5-224).
 * wtet_set_profiling resets the totals; kinds 0..12 are prepare_points,
 * vote, pick, split, split_points, flip_check (the FP32-only check),
 * flip_check_deferred (full predicates on the tets it deferred),
 * flip_decide_exec, flip_update, relocate, audit, output, flip_small (the
 * one-workgroup iterations of short lists).  Totals
 * accumulate over calls. */
wtet_status wtet_set_profiling(wtet_ctx *ctx, int on);
wtet_status wtet_kernel_times(const wtet_ctx *ctx, uint32_t kind, const char **name,
                              uint32_t *launches, double *total_ms);

/* Stage-level entry points: the loop of phase1::run (src/phase1.rs:16-1431)
 * driven one dispatch group at a time by the caller, as the reference's own
 * tests drive phase1::run and GpuState::dispatch_* (src/gpu/dispatch.rs).
 * One session per context, one cloud; any other triangulation call on the
 * context ends it.  The order within a round is vote, split, split_points,
 * then any number of flip iterations, then relocate before the next vote
 * (a wrong order returns WTET_ERR_INVALID_ARGUMENT).  Each call fills *info
 * (may be null) with the state after it. */
typedef struct wtet_stage_info {
  uint32_t round;      /* insertion rounds begun                             */
  uint32_t uninserted; /* points not yet inserted                            */
  uint32_t winners;    /* insertions of the current round                   */
  uint32_t tet_slots;  /* tet slots in use (alive or dead)                  */
  uint32_t active;     /* tets queued for the next flip check               */
  uint32_t checked;    /* tets checked by the last flip iteration           */
  uint32_t flips;      /* flips executed by the last flip iteration         */
  uint32_t violations; /* non-Delaunay faces found by the last gather       */
  uint32_t deferred;   /* tets the FP32 check of the last flip iteration
                          left to the full predicates                        */
} wtet_stage_info;

/* GpuState::new + sort + dispatch_init (src/gpu/mod.rs:48-204, :422-600,
 * src/gpu/dispatch.rs:33, init.wgsl:87-148): upload and normalize the
 * points, order them, five start tets, every point uninserted. */
wtet_status wtet_stage_init(wtet_ctx *ctx, const float *xyz, uint32_t n, const wtet_config *cfg,
                            wtet_stage_info *info);
/* dispatch_reset_votes + dispatch_vote + dispatch_pick_winner +
 * dispatch_build_insert_list (src/phase1.rs:113-166) and the compaction of
 * the uninserted list (dispatch_compact_vertex_arrays*, :1343-1352):
 * info->winners points are chosen, the rest stay uninserted. */
wtet_status wtet_stage_vote(wtet_ctx *ctx, wtet_stage_info *info);
/* dispatch_split (src/phase1.rs:305, split.wgsl:104-327): 1 -> 4 splits of
 * the winners' tets; the children are queued for the flip check. */
wtet_status wtet_stage_split(wtet_ctx *ctx, wtet_stage_info *info);
/* dispatch_split_points (src/phase1.rs:318, split_points.wgsl:222-335):
 * every uninserted point of a split tet moves into the child holding it. */
wtet_status wtet_stage_split_points(wtet_ctx *ctx, wtet_stage_info *info);
/* One flip iteration (src/phase1.rs:584-674 fast, :767-916 exact):
 * check the queued tets, vote, execute the accepted 2-3 / 3-2 flips,
 * rewire adjacency, queue the new tets.  predicate_mode -1: the session's;
 * 0 staged FP32 -> FP64 -> exact; 1 exact only (the decisions are the same). */
wtet_status wtet_stage_flip(wtet_ctx *ctx, int predicate_mode, wtet_stage_info *info);
/* dispatch_init_tet_to_flip + dispatch_update_flip_trace +
 * dispatch_relocate_points_fast (src/phase1.rs:1086-1104): uninserted
 * points of flipped tets to the tets now holding them. */
wtet_status wtet_stage_relocate(wtet_ctx *ctx, wtet_stage_info *info);
/* dispatch_gather (src/phase1.rs:1402, gather.wgsl:68-151): exact audit of
 * every finite face of the current mesh; info->violations. */
wtet_status wtet_stage_gather(wtet_ctx *ctx, wtet_stage_info *info);
/* GpuState::readback (src/gpu/mod.rs:207-332) of the current mesh: the
 * live tets in slot order as wtet_result rows (input ids, n..n+3 super,
 * n+4 infinity, (row << 5) | face adjacency); failed_verts from the last
 * gather if the mesh has not changed since.  No repair.  Free with
 * wtet_result_free. */
wtet_status wtet_stage_readback(wtet_ctx *ctx, wtet_result *out);
/* Where each uninserted point is: point_ids[i] (input id) lies in row
 * rows[i] of a readback of the same mesh state (rows follow slot order).
 * Valid after split_points or relocate; *count receives the number of
 * points (at most cap are written). */
wtet_status wtet_stage_locations(wtet_ctx *ctx, uint32_t cap, uint32_t *point_ids, uint32_t *rows,
                                 uint32_t *count);
/* The end of delaunay_3d (src/lib.rs:178-310) once every point is in:
 * remaining flip iterations, gather, readback, host repair; *out equals the
 * wtet_delaunay result for the same input.  Ends the session. */
wtet_status wtet_stage_finish(wtet_ctx *ctx, wtet_result *out);
/* Ends the session (no-op without one). */
wtet_status wtet_stage_end(wtet_ctx *ctx);

/* Host phase 2 alone (src/lib.rs:227-307: cpu_flip_violations ->
 * full_bw_reconstruction -> splay, src/phase2.rs): turns a valid
 * triangulation of the n normalized points (ids 0..n-1, super vertices
 * n..n+3, optional infinity tets with id n+4) into the exact Delaunay one.
 * Input ids that no tet uses are then inserted one Bowyer-Watson cavity at
 * a time (the mesh of the others must contain at least one tet of real or
 * super vertices), so the five start tets alone give the whole mesh.
 * *tets (num_tets*4) and *adjacency (same layout as wtet_result; may be
 * null: rebuilt by face matching) are read, then replaced by new malloc'd
 * arrays the caller frees with free().  regions (may be null) receives the
 * number of re-triangulated regions. */
wtet_status wtet_repair(const float *xyzn, uint32_t n, uint32_t **tets, uint32_t **adjacency,
                        uint32_t *num_tets, int mode, uint32_t *regions);

/* Insertion order of enable_hilbert_sorting (host copies of the device
 * code): the reference's first stride (calculate_initial_stride,
 * src/gpu/mod.rs:533-548), the position of each of n Hilbert-sorted points
 * in its stratified order (apply_strided_sampling_indexed, :573-600;
 * stride 0: the first stride for n), and the 3D Hilbert index of 8-bit
 * cells (xyz: n triples in 0..255).  The Hilbert curve is this library's
 * own (Skilling's method); the reference's comes from the hilbert_index
 * crate, which is not in the tree. */
uint32_t wtet_initial_stride(uint32_t n);
wtet_status wtet_stratified_ranks(uint32_t n, uint32_t stride, uint32_t *rank);
wtet_status wtet_hilbert_index(const uint32_t *xyz, uint32_t n, uint32_t *out);

/* Predicate batches (orient3d/insphere staging and SoS, src/predicates.rs).
 * kind: 0 orient3d_sos (4 pts, 4 ids), 1 orient4d (5 pts, 5 ids),
 *       2 orient3d raw sign (4 pts), 3 insphere raw sign (5 pts).
 * mode: 0 filtered, 1 exact only.  pts f32 xyz triples, one query after
 * another.  The _host variant runs the host copy used by the repair stage;
 * the _device variant runs the kernels' code on the GPU (host buffers in,
 * host buffers out) and adds stage counters to stats8 when non-null. */
wtet_status wtet_predicates_host(int kind, const float *pts, const uint32_t *ids, uint32_t n,
                                 int mode, int32_t *out);
wtet_status wtet_predicates_device(wtet_ctx *ctx, int kind, const float *pts,
                                   const uint32_t *ids, uint32_t n, int mode, int32_t *out,
                                   uint64_t *stats8);

#ifdef __cplusplus
}
#endif
#endif /* WTET_AMD_H */
