/*
 * grudon_oracle.h -- CPU oracle for the per-partition graph operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product (grudon_amd/), which must fail loudly when its HIP
 * library is missing instead of falling back to this code.
 *
 * It holds:
 *   1. the seeded graph generators (RMAT, Erdos-Renyi, edge weights) that the
 *      device generator must reproduce bit for bit;
 *   2. a restatement of the reference's partition layout
 *      (src/DistributedGraph.cpp:13-785, include/partitioning_engine/Partitioner.hpp:11-205);
 *   3. a restatement of the reference's BSP superstep
 *      (src/GraphAlgorithm.cpp:145-701) with the PR / CC / SSSP operators of
 *      src/graph_algorithms/{pr,cc,sssp}.cpp, including the bitvector +
 *      packed-value exchange of every (compute, memory) pair;
 *   4. closed-form oracles used to pin (3): BFS, Dijkstra, union-find
 *      connected components, FP64 power-iteration PageRank.
 *
 * Every function is pinned by a check that shares no code with it (golden
 * fixtures from textbook definitions, scipy / networkx routines, closed forms,
 * the reference's verify() invariants; the table per function is in
 * oracle/README.md), with one exception: the exchanged-byte count that the BSP
 * restatement reports in stats[1] is parity unpinned.
 */
#ifndef GRUDON_ORACLE_H
#define GRUDON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded generators (identical hash to the device generator) ---- */

/* SplitMix64 finaliser; the generators draw every random number from
 * orc_hash(seed, stream, index). */
uint64_t orc_hash(uint64_t seed, uint64_t index);

/* RMAT graph with 2^scale vertices and edge_factor * 2^scale raw edges,
 * quadrant probabilities a, b, c (d = 1 - a - b - c).  Self-loops and
 * duplicate edges are removed; symmetrize != 0 adds every reverse arc first.
 * Writes row_ptr[n+1] and col[<= cap]; returns the arc count, or -(needed)
 * when cap is too small. */
int64_t orc_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, int symmetrize,
                 uint32_t *row_ptr, uint32_t *col, int64_t cap);

/* Undirected Erdos-Renyi G(n, M) with M = n * mean_degree / 2 random vertex
 * pairs, symmetrized, without self-loops or duplicates. */
int64_t orc_er(uint32_t n, double mean_degree, uint64_t seed,
               uint32_t *row_ptr, uint32_t *col, int64_t cap);

/* Edge weights uniform in [1,255], a symmetric function of {u, v}. */
void orc_weights(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                 uint64_t wseed, uint32_t *w);

/* Build CSR (row_ptr[n+1], col) from raw (src, dst) pairs: sort, drop
 * self-loops and duplicates.  keys is scratch of length cnt (overwritten). */
int64_t orc_csr_from_pairs(uint32_t n, uint64_t *keys, int64_t cnt,
                           uint32_t *row_ptr, uint32_t *col);

/* ---- partition layout (reference DistributedGraph restatement) ---- */

typedef struct orc_part orc_part;

/* mirror may be NULL (naive gid % M).  M == 1 always uses the naive rule,
 * like the reference (Partitioner.hpp:73-85).  w may be NULL. */
orc_part *orc_part_build(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                         const uint32_t *w, int C, int M, const uint32_t *mirror);
void orc_part_free(orc_part *p);

/* Table queries used by the layout tests.
 *   what 0: T(c,m) translation-table size           (i=c, j=m)
 *   what 1: compute rTranslationTable(c,m) idx->gid  (out[T])
 *   what 2: memory  rTranslationTable(m,c) idx->gid  (i=m, j=c; out[T])
 *   what 3: memory partition m lid->gid             (i=m; out[nm])
 *   what 4: number of own vertices of memory m      (i=m)
 *   what 5: memory partition m local CSR row_ptr    (i=m; out[own+1])
 *   what 6: memory partition m local CSR col        (i=m; out[e_m])
 *   what 7: compute partition c size               (i=c)
 *   what 8: memory partition m size (own + cross)  (i=m)
 *   what 9: memory partition m edge count          (i=m)
 *   what 10: coverage by gid, 1 = covered           (out[n])
 * Returns the count (or the scalar) and fills out when non-NULL. */
int64_t orc_part_query(const orc_part *p, int what, int i, int j, uint32_t *out);

/* ---- BSP restatement ---- */

enum { ORC_PR = 0, ORC_CC = 1, ORC_SSSP = 2 };

/* stats[0] iterations, stats[1] exchange bytes (bitvectors + packed values,
 * both directions), stats[2] edges examined by gen_updates, stats[3] 1 if
 * MAX_ITERATIONS stopped the run. */

/* CC / SSSP on uint32 labels.  SSSP: weighted != 0 uses the edge weights,
 * else hop count (+1 per edge, src/graph_algorithms/sssp.cpp:69).  delta == 0
 * is the reference schedule (every improvement enters the next frontier);
 * delta > 0 enables the delta-stepping bucket rule.  out[gid]; trace (may be
 * NULL) receives the compute frontier size after every iteration. */
int orc_bsp_u32(orc_part *p, int algo, uint32_t source, int weighted,
                uint32_t delta, int max_iter, int nthreads,
                uint32_t *out, uint64_t *stats, uint32_t *trace);

/* Reference PageRank (delta push, src/graph_algorithms/pr.cpp) in FP32
 * (fp64 == 0) or FP64.  out[gid] = pr value. */
int orc_bsp_pr(orc_part *p, int fp64, int max_iter, int nthreads,
               double *out, uint64_t *stats, uint32_t *trace);

/* The same runs with a superstep schedule (src/GraphAlgorithm.cpp:169-195):
 *   sched 0: the offloaded superstep every iteration (NDP_OFFLOAD; what the
 *            reference runs, its vote being forced at :172-174);
 *   sched 1: the NDPEngine vote (include/offload_engine/NDPEngine.hpp:10-51)
 *            on every compute partition, combined by MIN: NO_OFFLOAD (edge
 *            shipping, :426-474 and :582-637) when every partition's frontier
 *            is non-empty, at most n/20 and covered;
 *   sched 2: edge shipping whenever the whole frontier is covered.
 * A shipping iteration sends [gid, out-neighbour gids] (4 bytes each, plus
 * 4 bytes of weight per arc for weighted SSSP) from the memory partitions to
 * the masters and has no reduce; the master aggregates with the algorithm's
 * own operator.  votes (may be NULL, max_iter entries) receives 1 for every
 * iteration that shipped. */
int orc_bsp_u32_sched(orc_part *p, int algo, uint32_t source, int weighted,
                      uint32_t delta, int max_iter, int nthreads, int sched,
                      uint32_t *out, uint64_t *stats, uint32_t *trace, uint8_t *votes);
int orc_bsp_pr_sched(orc_part *p, int fp64, int max_iter, int nthreads, int sched,
                     double *out, uint64_t *stats, uint32_t *trace, uint8_t *votes);

/* One pair's exchange at iteration `iter` of the offloaded CC / SSSP run:
 * dir 0 = the broadcast bitvector c -> m with its packed properties
 * (GraphAlgorithm.cpp:199-274), dir 1 = the reduce bitvector m -> c with its
 * packed updates (:348-425).  Fills gids (ascending) of the set bits and the
 * value sent for each; returns their number, or < 0 for a bad argument
 * (0 also when the run ended before `iter`). */
int64_t orc_bsp_pair_exchange(orc_part *p, int algo, uint32_t source, int weighted,
                              uint32_t delta, int iter, int dir, int c, int m,
                              uint32_t *gids, uint32_t *vals);

/* ---- closed forms ---- */

void orc_bfs(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
             uint32_t src, uint32_t *dist);
void orc_dijkstra(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                  const uint32_t *w, uint32_t src, uint32_t *dist);
/* min gid per weakly connected component */
void orc_cc_unionfind(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                      uint32_t *label);
/* Normalised PageRank with uniform dangling redistribution, FP64:
 *   r' = (1-d)/n + d * (sum_{u->v} r_u/deg_u + D/n),  D = sum_{deg_u=0} r_u,
 * r0 = 1/n.  Runs exactly `iters` iterations when iters > 0, else until the
 * L1 change drops below tol (cap 10000).  Returns the iteration count;
 * *last_residual = L1 change of the last iteration. */
int orc_pr_pull(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                double damping, int iters, double tol, double *r,
                double *last_residual);
/* Same recurrence, evaluated as a pull over the transposed graph with
 * nthreads threads (per-vertex FP64 sums in ascending source order). */
int orc_pr_pull_mt(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                   double damping, int iters, double tol, double *r,
                   double *last_residual, int nthreads);

/* ---- full-size checks and the timed CPU pull (bench.py) ---- */

/* The threaded FP64 pull of orc_pr_pull_mt with its transposed graph built
 * once (in parallel; sources ascending per row, as orc_pr_pull_mt sums
 * them), so that solves can be repeated and timed: orc_pull_plan_run starts
 * from r0 = 1/n and sets *seconds to the wall time of the iteration loop. */
typedef struct orc_pull_plan orc_pull_plan;
orc_pull_plan *orc_pull_plan_build(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, int nthreads);
int orc_pull_plan_run(orc_pull_plan *p, double damping, int iters, double tol, double *r,
                      double *last_residual, double *seconds);
uint64_t orc_pull_plan_arcs(const orc_pull_plan *p);
void orc_pull_plan_free(orc_pull_plan *p);
/* min gid per weakly connected component by lock-free parallel union-find */
void orc_cc_unionfind_mt(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, uint32_t *label,
                         int nthreads);
/* Number of violated Bellman conditions of dist from src (0: dist holds the
 * shortest-path distances; w NULL = +1 per arc; unreached = 2^32 - 1). */
uint64_t orc_sssp_check_mt(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, const uint32_t *w,
                           const uint32_t *dist, uint32_t src, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
