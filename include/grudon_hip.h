/*
 * grudon_hip.h -- C-ABI of the MI355X (gfx950) graph operator library
 * (libgrudon_hip.so).  Plain pointers and sizes only; every function returns
 * a grd_status and never throws.  grd_last_error() gives the message of the
 * calling thread's last failure.
 *
 * The reference drives its per-partition operator path through in-process
 * C++ virtual calls (include/GraphAlgorithm.hpp:24-86, include/Workers.hpp:12-115)
 * and MPI point-to-point messages (include/MPI.hpp:8-124).  Here one process
 * owns every compute and memory partition on one GPU; the exchange between
 * partitions is done by device kernels over HBM-resident wire buffers.  The
 * entry points below are what those virtual calls bind to in bulk form.
 *
 * Threading: one host thread per grd_ctx.  All operations are ordered on the
 * context's HIP stream.
 */
#ifndef GRUDON_HIP_H
#define GRUDON_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GRD_OK = 0,
  GRD_ERR_ARG = 1,    /* invalid argument or size beyond a stated limit */
  GRD_ERR_HIP = 2,    /* HIP runtime error (message in grd_last_error) */
  GRD_ERR_OOM = 3,    /* device or host allocation failed */
  GRD_ERR_STATE = 4,  /* call out of order (e.g. step before init) */
  GRD_ERR_IO = 5,     /* graph / partition / weight file could not be read */
} grd_status;

typedef struct grd_ctx grd_ctx;
typedef struct grd_graph grd_graph;
typedef struct grd_state grd_state;

/* Algorithms.  GRD_PR is the reference's delta-push PageRank
 * (src/graph_algorithms/pr.cpp), GRD_CC src/graph_algorithms/cc.cpp,
 * GRD_SSSP src/graph_algorithms/sssp.cpp (hop count, or edge weights and
 * delta-stepping as extensions), GRD_PR_PULL the normalised pull PageRank
 * (FP32 Kahan contributions, FP64 dangling mass and residual). */
typedef enum { GRD_PR = 0, GRD_CC = 1, GRD_SSSP = 2, GRD_PR_PULL = 3 } grd_algo;

/* Per-iteration schedule option.  0-3 are the offloaded superstep
 * (NDP_OFFLOAD, src/GraphAlgorithm.cpp:199-581): bit 0 = pull traversal
 * (else push), bit 1 = dense sync (else sparse bitvector + packed values).
 * GRD_OPT_SHIP_EDGES is the NO_OFFLOAD superstep (GraphAlgorithm.cpp:426-474,
 * 582-637): after the sparse broadcast, the memory partitions ship the
 * frontier's edge lists to the masters, which apply the updates
 * themselves; there is no reduce.  It needs every frontier vertex covered
 * (all out-neighbours on the same master, see grd_graph_coverage). */
enum {
  GRD_OPT_PUSH_SPARSE = 0,
  GRD_OPT_PULL_SPARSE = 1,
  GRD_OPT_PUSH_DENSE = 2,
  GRD_OPT_PULL_DENSE = 3,
  GRD_OPT_SHIP_EDGES = 4
};

/* Schedule selection policy for grd_run. */
typedef enum {
  GRD_SCHED_PLAIN = 0,     /* the reference schedule: push + sparse every iteration */
  GRD_SCHED_HEURISTIC = 1, /* direction-optimizing heuristic */
  GRD_SCHED_LEARNED = 2,   /* MLP selector with the heuristic as fallback */
  GRD_SCHED_FIXED = 3,     /* the option in grd_params.fixed_option every iteration
                              (GRD_OPT_SHIP_EDGES falls back to the offloaded
                              push while the frontier is not covered) */
  GRD_SCHED_NDP_ENGINE = 4 /* the reference's offload vote (NDPEngine,
                              include/offload_engine/NDPEngine.hpp:10-51, which
                              run() overrides at GraphAlgorithm.cpp:172-174):
                              edge shipping when every compute partition's
                              frontier is non-empty, at most n/20 and covered,
                              else the offloaded push */
} grd_sched;

typedef struct {
  uint32_t source;       /* SSSP source gid (reference: gid 0, sssp.cpp:39-46) */
  uint32_t weighted;     /* SSSP: use edge weights (else +1 per edge) */
  uint32_t delta;        /* SSSP delta-stepping bucket width; 0 = reference schedule */
  uint32_t max_iter;     /* reference MAX_ITERATIONS = 1000 (GraphAlgorithm.hpp:10) */
  double damping;        /* PR_PULL damping (0.85) */
  double tolerance;      /* PR_PULL convergence bound on the L1 change per iteration */
  uint32_t sched;        /* grd_sched */
  uint32_t fixed_option; /* option for GRD_SCHED_FIXED */
  uint32_t time_iters;   /* record per-iteration device time in the trace */
  uint32_t time_kernel;  /* bracket the dominant traversal kernel with HIP events;
                            2 = the events only: a CC / SSSP pull then counts
                            the gathers of rows that gathered none as made
                            (1 counts them exactly with a pass of its own,
                            outside the events but inside the solve's time) */
  uint32_t offload_only; /* heuristic / learned: never turn a push superstep into
                            edge shipping (by default they ship whenever the
                            frontier is covered, the cheaper push on one GPU) */
  uint32_t pull_every_row; /* CC / SSSP pull supersteps: gather every row's arcs.
                              By default a row whose vertex already holds a value
                              at or below the smallest value broadcast this
                              superstep is not pulled (no candidate can improve
                              it); results and iteration counts are the same
                              either way */
} grd_params;

/* One record per iteration (grd_run trace).  feat holds the selector's 16
 * input features for the iteration. */
typedef struct {
  uint32_t iteration;
  uint32_t option;          /* schedule option that ran */
  uint64_t frontier;        /* vertices broadcast this iteration */
  uint64_t frontier_edges;  /* their out-degree sum */
  uint64_t dirty;           /* memory proxies sent back (reduce) */
  uint64_t next_frontier;   /* frontier after the update */
  double ms;                /* device time of the iteration (time_iters) */
  float feat[16];
} grd_iter_rec;

typedef struct {
  uint64_t iterations;
  uint64_t bytes_exchanged;  /* bitvectors + packed values, both directions */
  uint64_t edges_examined;
  uint64_t hit_max_iter;
  double ms;                 /* device time of the whole iteration loop */
  double residual;           /* PR_PULL: last L1 change */
  uint64_t option_counts[5]; /* supersteps run with each GRD_OPT_* */
  uint64_t fallbacks;        /* learned: iterations decided by the heuristic;
                                fixed edge shipping: iterations run offloaded */
  /* time_kernel: the dominant traversal kernel (pull: k_pull / k_prp_pull,
   * push: k_push, edge shipping: k_ship_send + k_ship_recv), summed over
   * launches, with its algorithmic HBM bytes: pull 8 B per edge + 8 B per
   * row (a CC / SSSP pull: of the arcs and rows it reads, the weight as
   * stored, the rows-done bitmap), push 12 B per edge + 16 B per vertex,
   * shipping 24 B per edge (36 weighted) + 36 B per record */
  double kernel_ms;
  uint64_t kernel_launches;
  uint64_t kernel_bytes;
} grd_run_stats;

/* ---- context: replaces MPICore (include/MPI.hpp:11-24) ---- */
const char *grd_last_error(void);
grd_status grd_create(int device, grd_ctx **out);
grd_status grd_destroy(grd_ctx *ctx);
/* Host threads used for graph construction (reference -t, src/main.cpp:25). */
grd_status grd_set_host_threads(grd_ctx *ctx, int nthreads);
/* Where the partitioned layout (DistributedGraph ctor,
 * src/DistributedGraph.cpp:13-785) is built: on_device != 0 (the default)
 * with device kernels, 0 with host threads.  Both give the same layout. */
grd_status grd_set_layout(grd_ctx *ctx, int on_device);

/* No counterpart in the reference: how pull-PageRank states (GRD_PR_PULL)
 * created afterwards run their pull.
 *   mode 0: the row-block pull over the CSC;
 *   mode 1: LDS tiles -- the CSC regrouped by (memory partition, source
 *     tier, row), a tier being 2^tier_bits consecutive proxies (4..15, 0 =
 *     15), every gather served from the tier's window staged in LDS by
 *     tiles of about tile_arcs arcs (0 = 65536);
 *   mode 2: L2 segments -- the same regrouping with tiers of 2^tier_bits
 *     proxies (4..24, 0 = 19: 2 MB of values, resident in L2 while their row
 *     blocks run) and the row-block pull's kernel over them;
 *   mode 3: hot / cold split -- each row's arcs from its partition's
 *     2^tier_bits hottest sources (4..15, 0 = 15) summed from an LDS window
 *     into a per-row partial, the rest of the row pulled by the row-block
 *     kernel, which merges that partial first (tile_arcs must be 0); the
 *     default, and the fastest measured.  Mode 3 also serves the pull
 *     supersteps of CC, SSSP and reference PageRank (states of any algorithm
 *     created afterwards); modes 1 and 2 serve GRD_PR_PULL only.
 * Modes 1-3 regroup a graph when its first pull-PageRank state is created,
 * and a graph keeps the regrouping it was first given; per row, the
 * partials are merged in a fixed order.  All modes compute the same sums
 * within FP32 rounding. */
grd_status grd_set_pull_tiling(grd_ctx *ctx, int mode, uint32_t tier_bits, uint32_t tile_arcs);

/* ---- graphs ---- */

/* Seeded device generators; fill caller host arrays with the global CSR
 * (row_ptr[n+1], col[cap]); *m_out = arc count.  Self-loops and duplicates
 * are removed.  RMAT: n = 2^scale, edge_factor * n raw edges. */
grd_status grd_generate_rmat(grd_ctx *ctx, int scale, int edge_factor, double a, double b, double c,
                             uint64_t seed, int symmetrize, uint32_t *row_ptr, uint32_t *col,
                             uint64_t cap, uint64_t *m_out);
grd_status grd_generate_er(grd_ctx *ctx, uint32_t n, double mean_degree, uint64_t seed, uint32_t *row_ptr,
                           uint32_t *col, uint64_t cap, uint64_t *m_out);

/* Seeded edge weights uniform in [1,255] for a host CSR: a symmetric function
 * of the endpoints {u, v}, hash(wseed, min << 32 | max). */
grd_status grd_edge_weights(grd_ctx *ctx, uint32_t n, const uint32_t *row_ptr, const uint32_t *col, uint64_t wseed,
                            uint32_t *w);

/* Partition a global CSR into C compute and M memory partitions resident in
 * HBM (reference DistributedGraph ctor, src/DistributedGraph.cpp:13-785;
 * master = gid % C, mirror = gid % M or mirror[gid], Partitioner.hpp:11-205).
 * weights may be NULL; mirror may be NULL (naive). */
grd_status grd_graph_from_csr(grd_ctx *ctx, uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                              const uint32_t *weights, int num_compute, int num_memory,
                              const uint32_t *mirror, grd_graph **out);
/* Generate on the device and partition in one call.  wseed != 0 attaches
 * weights uniform in [1,255] (a symmetric function of the endpoints). */
grd_status grd_graph_rmat(grd_ctx *ctx, int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                          int symmetrize, uint64_t wseed, int num_compute, int num_memory, grd_graph **out);
grd_status grd_graph_er(grd_ctx *ctx, uint32_t n, double mean_degree, uint64_t seed, uint64_t wseed,
                        int num_compute, int num_memory, grd_graph **out);
/* Read a Galois .gr (v1/v2) or MatrixMarket .mtx file; partition_file (may
 * be NULL/empty) holds one memory-partition id per line (Partitioner.hpp:89-105). */
grd_status grd_graph_load(grd_ctx *ctx, const char *path, const char *partition_file, int num_compute,
                          int num_memory, grd_graph **out);
grd_status grd_graph_free(grd_graph *g);
/* Read a .gr / .mtx file into caller buffers: call once with row_ptr NULL to
 * get *n and *m, then again with row_ptr[n+1], col[m] (and weights[m], may be
 * NULL) allocated.  Duplicates and self-loops are kept as in the file. */
grd_status grd_graph_read(const char *path, uint32_t *n, uint64_t *m, uint32_t *row_ptr, uint32_t *col,
                          uint32_t *weights, int *has_weights);
/* Write a Galois .gr version 1 file (u32 edge data when weights != NULL);
 * replaces the graph-convert step of scripts/convert_and_partition.sh. */
grd_status grd_graph_write_gr(const char *path, uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                              const uint32_t *weights);
/* info[0] n, [1] m, [2] C, [3] M, [4] compute slots, [5] memory proxies,
 * [6] cross mirrors, [7] weighted, [8] device bytes held by the graph,
 * [9] pull row blocks, [10] long rows; [11]-[14] describe the pull mode's
 * reorganisation (0 until a state of the graph built it): modes 1 and 2
 * [11] (row, tier) pairs, [12] tiles, [13] row blocks, [14] long pairs;
 * mode 3 [11] rows (one partial each), [12] hot workgroup ranges, [13] hot
 * wave blocks, [14] long hot rows. */
grd_status grd_graph_info(const grd_graph *g, uint64_t *info16);
/* Translation table of pair (c, m) as gids in wire-index order: the
 * reference sTranslationTable/rTranslationTable content
 * (DistributedGraph.cpp:701-734).  Returns T in *count; out may be NULL. */
grd_status grd_graph_table(const grd_graph *g, int c, int m, uint32_t *out_gids, uint64_t *count);
/* Out-degree per gid (host copy). */
grd_status grd_graph_out_degree(const grd_graph *g, uint32_t *out);
/* Coverage vector (reference coverage_vector, src/DistributedGraph.cpp:286-370,
 * read by NDPEngine.hpp:23-26): covered[gid] = 1 when every out-neighbour of
 * gid has gid's master (gid % C), so that edge shipping can apply all of its
 * updates on that compute partition.  *n_uncovered = vertices not covered.
 * Either output may be NULL. */
grd_status grd_graph_coverage(const grd_graph *g, uint8_t *covered, uint64_t *n_uncovered);
/* No counterpart in the reference: 16 FNV-1a digests of the partitioned
 * layout (device arrays, host tables, sizes), so that two builds of the same
 * graph can be compared (e.g. the device and the host construction). */
grd_status grd_graph_digest(const grd_graph *g, uint64_t *digest16);

/* ---- algorithm state: GraphAlgorithm<VP> (GraphAlgorithm.hpp:24-86) ---- */
/* A state with 4-byte vertex properties: GraphAlgorithm<float> (GRD_PR,
 * GRD_PR_PULL) or GraphAlgorithm<uint32_t> (GRD_CC, GRD_SSSP). */
grd_status grd_state_create(grd_graph *g, int algo, grd_state **out);
/* A state for the property width value_bytes of the reference's
 * instantiation (GraphAlgorithm.hpp:88-92: float, double, uint64_t,
 * uint32_t; CC<uint64_t> cc.hpp:32, SSSP<uint64_t> sssp.hpp:33,
 * PageRank<double> pr.hpp:39).  value_bytes 4 is grd_state_create.
 * value_bytes 8 is accepted for GRD_CC and GRD_SSSP: labels and distances
 * are gids and path lengths below 2^32 - 1 (weighted SSSP needs
 * 255 * (n - 1) < 2^32 - 1), so the device computes them in 32 bits and
 * grd_download / grd_pair_exchange return them widened to uint64 (the
 * unreached value 2^32 - 1 becomes 2^64 - 1, numeric_limits<uint64_t>::max,
 * sssp.cpp:25); the exchanged bytes count 8 per value, as the reference's
 * MPI datatype would.  value_bytes 8 for GRD_PR / GRD_PR_PULL
 * (PageRank<double>) returns GRD_ERR_ARG: FP64 vertex properties are not
 * built (the pull form keeps FP64 for the dangling mass and the residual). */
grd_status grd_state_create_typed(grd_graph *g, int algo, uint32_t value_bytes, grd_state **out);
grd_status grd_state_free(grd_state *s);
/* init(): pr.cpp:20-61 / cc.cpp:18-40 / sssp.cpp:18-47 */
grd_status grd_init(grd_state *s, const grd_params *p);
/* The superstep in bulk form (GraphAlgorithm.cpp:177-690):
 *   grd_broadcast  compute -> memory: frontier bitvector + values   (:199-343)
 *   grd_traverse   gen_updates on every memory partition           (:348-356)
 *   grd_reduce     memory -> compute: dirty bitvector + values and
 *                  aggregate() into vertex_updates                  (:377-581)
 *   grd_update     update_frontier + apply_updates + termination    (:643-662)
 * option = GRD_OPT_*; *frontier_out = next frontier size. */
grd_status grd_broadcast(grd_state *s, int option);
grd_status grd_traverse(grd_state *s, int option);
grd_status grd_reduce(grd_state *s, int option);
grd_status grd_update(grd_state *s, uint64_t *frontier_out);
/* Per-element aggregate(lid, buffer_val) of the reference API
 * (GraphAlgorithm.hpp:42, dispatched by UpdateWorker::aggregate,
 * src/Workers.cpp:71-75): combine value into vertex_updates of the vertex
 * with local id lid on compute partition compute_partition, as run() calls it
 * with rTranslationTable[m][bit] for every set bit of a reduce message
 * (src/GraphAlgorithm.cpp:565-578).  Compute local ids follow the reference
 * (DistributedGraph.cpp:170-191): lid = gid / C on partition gid % C.  PR
 * adds (addUpdate), CC and SSSP take the minimum (minUpdate); the vertex is
 * marked updated.  value points to a float (PR), a uint32 or, for a state
 * created with value_bytes 8, a uint64.  Safe to call from one host thread;
 * the bulk form is grd_reduce. */
grd_status grd_aggregate_one(grd_state *s, int compute_partition, uint32_t lid, const void *value);
/* Whole loop with the chosen schedule policy; trace (may be NULL) receives
 * min(iterations, trace_cap) records. */
grd_status grd_run(grd_state *s, grd_run_stats *stats, grd_iter_rec *trace, uint32_t trace_cap);
/* The schedule option the state's policy picks for the next superstep (the
 * offload decision point of GraphAlgorithm.cpp:187-195). */
grd_status grd_select_option(grd_state *s, int *option);
/* Counters of the iterations run so far (step-wise use). */
grd_status grd_state_stats(const grd_state *s, grd_run_stats *out);
/* Result by gid: CC labels / SSSP distances (uint32, or uint64 for a state
 * created with value_bytes 8), PR values (float). */
grd_status grd_download(grd_state *s, void *out_by_gid);
/* The current superstep's exchange of one pair, as the reference's public
 * Worker::bitCommVector and the packed propertyBuffers hold it
 * (include/Workers.hpp:30-45; filled at GraphAlgorithm.cpp:199-274 and
 * :348-425): dir 0 = the broadcast c -> m (after grd_broadcast), dir 1 = the
 * reduce m -> c (after grd_reduce).  gids[] receives the vertices whose bit
 * is set, ascending (the device's bit order is the degree order), values[]
 * the property or update packed for each (4 bytes, or 8 for a state created
 * with value_bytes 8); *count their number.
 * gids and values may be NULL.  GRD_ERR_STATE when that exchange has not
 * run in this superstep (an edge-shipping superstep has no reduce). */
grd_status grd_pair_exchange(grd_state *s, int dir, int c, int m, uint32_t *gids, void *values, uint64_t *count);
/* Selector features of the current state (16 floats). */
grd_status grd_features(grd_state *s, float *out16);
/* Copy / restore every mutable array of the state (used for labelling). */
grd_status grd_snapshot(grd_state *s);
grd_status grd_restore(grd_state *s);

/* ---- schedule selector (learned MLP 16 -> 64 -> 64 -> 4) ---- */
/* Load BF16 weights (5,508 values: W1[64][16] b1[64] W2[64][64] b2[64]
 * W3[4][64] b3[4]) into the
 * context; the learned policy falls back to the heuristic when none is loaded.
 * The network overrules the heuristic only where its top logit exceeds the
 * runner-up by at least min_margin. */
grd_status grd_selector_load(grd_ctx *ctx, const uint16_t *bf16_weights, uint64_t count, float min_margin);
/* The margin to load the packaged weights (grudon_amd/selector_bf16.bin)
 * with: the one at which the learned schedule was fastest against the
 * heuristic on graphs outside the training draw (RMAT-22 seed 9090,
 * Erdos-Renyi 2^20 seed 8080; PageRank, CC, hop and weighted SSSP). */
#define GRD_SELECTOR_MARGIN 1.0f
/* Train on the GPU with the project's forward/backward kernels: feat[n*16],
 * label[n] in 0..3, cost[n*4] (per-option times, used to weight the loss).
 * Writes the trained BF16 weights to out (may be NULL) and loads them. */
grd_status grd_selector_train(grd_ctx *ctx, const float *feat, const uint32_t *label, const float *cost,
                              uint64_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                              uint16_t *out_bf16, float *final_loss);
/* Forward pass on the GPU: logits[n*4]. */
grd_status grd_selector_infer(grd_ctx *ctx, const float *feat, uint64_t n, float *logits);
/* No counterpart in the reference; for callers that cannot allocate device
 * memory (tight serving loops, code that forbids hipMalloc): the same
 * forward pass as grd_selector_infer, bit for bit, with no allocation.  The
 * features and logits are staged in a fixed device array of the library
 * (320 KB per GPU, held until the process ends), so n is at most 4096
 * (GRD_ERR_ARG beyond).  Calls that overlap on one GPU share that array
 * and race; the caller still passes the host inputs and outputs. */
grd_status grd_selector_infer_noalloc(grd_ctx *ctx, const float *feat, uint64_t n, float *logits);
/* One forward+backward on the GPU from FP32 weights w[5508] (no BF16
 * rounding); sample_weight[n] may be NULL (all 1).  Writes the summed
 * weighted cross-entropy and its gradient (for the CPU gradient check). */
grd_status grd_selector_grad(grd_ctx *ctx, const float *w, const float *feat, const uint32_t *label,
                             const float *sample_weight, uint64_t n, float *loss, float *grad);

/* ---- batched form: many graphs in one persistent launch ---- */
typedef struct grd_batch grd_batch;
grd_status grd_batch_create(grd_ctx *ctx, uint32_t count, const uint32_t *n, const uint64_t *m,
                            const uint32_t *const *row_ptr, const uint32_t *const *col, grd_batch **out);
grd_status grd_batch_free(grd_batch *b);
/* Tunables (0 = keep): tile = arcs per row block and long-row threshold
 * (1024, 2048 or 4096), grid_per_cu = persistent workgroups per CU (<= 64),
 * K = iterations per captured hipGraph (<= 64). */
grd_status grd_batch_tune(grd_batch *b, uint32_t tile, uint32_t grid_per_cu, uint32_t K);
/* Autotune the tunables by timing candidate settings on this batch
 * (tile x grid_per_cu x K); chosen3 (may be NULL) receives the winner. */
grd_status grd_batch_autotune(grd_batch *b, uint32_t *chosen3);
/* PageRank pull (Kahan FP32, FP64 dangling/residual) on every graph, with
 * iterations captured in a hipGraph; iters_out[count], ms_out may be NULL. */
grd_status grd_batch_pagerank(grd_batch *b, double damping, double tolerance, uint32_t max_iter,
                              uint32_t *iters_out, double *ms_out);
grd_status grd_batch_cc(grd_batch *b, uint32_t max_iter, uint32_t *iters_out, double *ms_out);
/* Results of graph i: PR (float[n]) or CC labels (uint32[n]). */
grd_status grd_batch_download(grd_batch *b, uint32_t i, int algo, void *out);

#ifdef __cplusplus
}
#endif
#endif
