// grd_internal.hpp -- host-side structures of libgrudon_hip.so.
//
// Device layout (all partitions of one graph resident in one GPU's HBM):
//
//   compute slots  [0, S): compute partition c, then mirror group m, each
//                  (c, m) group padded to a multiple of 64 slots.  A group
//                  holds the "own-on-m" vertices of the reference's compute
//                  sTranslationTable[m] (DistributedGraph.cpp:701-717),
//                  ordered by out-degree, highest first, ties by ascending
//                  gid; the wire index of a slot is slot - cbase(c, m).
//   memory proxies [0, P): memory partition m, then master c, then the own
//                  region (same order as the compute group, so bitvector
//                  words map one to one) and the cross-mirror region
//                  (ordered by the vertices' compute slots), each padded to 64.  The (m, c) own+cross
//                  range holds the reference's memory-side translation table
//                  for master c (DistributedGraph.cpp:644-660).
//
// Both orders are permutations of the reference's wire order inside each
// bitvector (same length, same popcount, same bytes moved).  The pull gathers
// a source's value once per out-arc, so it packs the most gathered values of a
// memory partition into the fewest cache lines.
//
// Because every pair region starts on a 64-bit word, a pair's bitvector is a
// word-aligned slice of the frontier / dirty bitmaps; bitvector sync copies
// words and packs values by popcount prefix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <array>
#include <functional>
#include <string>
#include <vector>

#include "../../include/grudon_hip.h"

namespace grd {

void set_error(const std::string &msg);
grd_status fail(grd_status st, const std::string &msg);

#define GRD_HIP(expr)                                                                            \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      return ::grd::fail(GRD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));       \
  } while (0)

#define GRD_TRY(expr)                 \
  do {                                \
    grd_status _s = (expr);           \
    if (_s != GRD_OK) return _s;      \
  } while (0)

constexpr uint32_t kNone = 0xFFFFFFFFu;

inline uint64_t pad64(uint64_t x) { return (x + 63) & ~uint64_t(63); }

// Device allocation list owned by a graph or a state.
struct DevMem {
  std::vector<void *> ptrs;
  uint64_t bytes = 0;
  grd_status alloc(void **p, uint64_t nbytes);
  template <typename T>
  grd_status alloc_n(T **p, uint64_t n) { return alloc(reinterpret_cast<void **>(p), n * sizeof(T)); }
  void release();
  ~DevMem() { release(); }
};

// Counters written by device kernels and read back once per iteration.
struct Counters {
  unsigned long long frontier;        // popcount of the new frontier
  unsigned long long frontier_edges;  // out-degree sum of the new frontier
  unsigned long long dirty;           // dirty memory proxies at reduce
  unsigned long long pending;         // SSSP delta-stepping: pending vertices
  unsigned long long improved;        // compute vertices improved by update
  unsigned long long heavy;           // push: heavy-vertex chunks appended this iteration
  unsigned long long edges;           // edges examined by the traversal
  unsigned int frontier_maxdeg;
  unsigned int pend_min;
  double residual;                    // PR pull: L1 change
  double dangling[2];                 // PR pull: FP64 dangling mass (ping-pong)
  unsigned long long heavy_edges;     // push: edges of heavy vertices (k_push_heavy)
  unsigned long long frontier_uncovered;  // frontier vertices with an out-neighbour of another master
  unsigned long long ship_error;      // edge shipping: the active list disagreed with the sized wire
  unsigned int fmin;                  // CC / SSSP pull: smallest value broadcast this superstep
  unsigned int pad32;
  unsigned long long pull_skipped;    // CC / SSSP pull: arcs of the cold row blocks left out (not read)
  unsigned long long pull_skipped_hot;   // and of the hot wave blocks left out
  unsigned long long pull_skipped_rows;  // rows of the cold row blocks left out
  unsigned long long pull_done_arcs;     // cold arcs of rows that gathered none (measured with time_kernel)
  float sel[8];                       // learned selector: the next superstep's 4 logits and argmax
  unsigned long long fpart[64];       // frontier size of each compute partition (C > 1)
};

// pull row block: rows [r0, r0 + nrows) with arcs [e0, e1); nrows == 0 marks
// one segment [e0, e1) of the long row r0
struct PullBlock {
  uint32_t r0, nrows, e0, e1;
};
struct PullLong {
  uint32_t row, first, nseg, pad;
};
// tiled PageRank pull: one tile = a run of row blocks of one (memory
// partition, source tier) group, whose tier window mprop[wbase, wbase + wlen)
// is staged in LDS once
struct TierTile {
  uint32_t b0, nb, wbase, wlen;
};
constexpr uint32_t kPullNnz = 2048;     // arcs per block (8 per thread)
constexpr uint32_t kPullMaxRows = 1024;  // rows per block
constexpr uint32_t kPullHot = 2048;      // PR pull: hottest sources of a partition kept in LDS
constexpr uint32_t kHeavyDeg = 4096;     // push: out-degree handled by k_push_heavy, in chunks of this many arcs
constexpr uint32_t kShipTile = 1024;     // edge shipping: wire positions per workgroup tile
constexpr int kMaxParts = 64;            // largest C and M
constexpr uint32_t kTierBitsMax = 15;    // tiled pull: largest tier, 2^15 sources (128 KB of LDS)
constexpr uint32_t kTierTileArcs = 65536;  // tiled pull: default arcs per tile
constexpr uint32_t kSegBitsMax = 24;       // segmented pull: largest segment, 2^24 sources
constexpr uint32_t kSegBits = 19;          // segmented pull: default segment, 2^19 sources (2 MB of values)
constexpr uint32_t kHotWin = 32768;        // hot / cold pull: hottest sources of a partition in LDS (128 KB)

}  // namespace grd

struct grd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int host_threads = 0;
  int num_cus = 256;
  int pull_mode = 3;        // pull PageRank: 0 row-block pull, 1 LDS-tiled, 2 L2-segmented, 3 hot / cold
                            // split (the default; tiers.hip)
  uint32_t tier_bits = 0;   // tunables for graphs regrouped later (0 = the mode's default)
  uint32_t tile_arcs = 0;
  bool layout_on_device = true;  // build the partitioned layout with device kernels (else on the host)
  // learned selector weights (BF16, device) and threshold
  uint16_t *sel_w = nullptr;
  bool sel_loaded = false;
  float sel_margin = 0.0f;
  float *sel_out = nullptr;       // device: 4 logits + chosen option
  float *sel_out_host = nullptr;  // pinned
};

struct grd_graph {
  grd_ctx *ctx = nullptr;
  uint32_t n = 0;
  uint64_t m = 0;
  int C = 1, M = 1;
  bool weighted = false;
  uint32_t S = 0;       // compute slots (padded)
  uint32_t P = 0;       // memory proxies (padded)
  uint64_t X = 0;       // cross mirrors
  // host metadata
  std::vector<uint32_t> cbase, ccnt;            // [C*M] pair (c,m): slot base and own count
  std::vector<uint32_t> obase, ocnt, xbase, xcnt;  // [M*C] pair (m,c): proxy bases and counts
  std::vector<uint32_t> h_gid_slot;             // gid -> compute slot
  std::vector<uint32_t> h_deg;                  // gid -> out-degree
  std::vector<uint32_t> h_proxy_gid;            // proxy -> gid (kNone for padding)
  double avg_degree = 0.0;
  uint64_t n_sinks = 0;   // vertices with out-degree 0
  uint32_t max_deg = 0;   // largest out-degree
  // device arrays
  grd::DevMem mem;
  uint32_t *slot_gid = nullptr;    // [S]
  uint32_t *slot_deg = nullptr;    // [S] global out-degree
  uint32_t *word_dst = nullptr;    // [S/64] proxy word receiving each slot word
  uint32_t *proxy_slot = nullptr;  // [P]
  uint32_t *rp = nullptr, *col = nullptr, *w = nullptr;     // push CSR over proxies
  uint32_t *crp = nullptr, *ccol = nullptr, *cw = nullptr;  // pull CSC over proxies
  uint32_t *xptr = nullptr, *xlist = nullptr;               // [S+1], [X] cross proxies per slot
  // edge shipping (NO_OFFLOAD schedule)
  uint8_t *word_c = nullptr;     // [S/64] compute partition of each slot word
  uint8_t *word_nvalid = nullptr;  // [S/64] valid slots at the start of each slot word (PR pull; built on demand)
  // [P/64] rows with at least one arc in the CSC (CC / SSSP pull: a row block
  // is left out when its rows with arcs are done; built on first use).  The
  // whole CSC counts for the cold blocks of the hot / cold split too: a row
  // without cold arcs still takes its hot partial in the cold pass.
  uint64_t *rarc = nullptr;
  uint32_t wmin = 0;         // smallest arc weight (weighted SSSP pull; found on first use)
  bool wmin_known = false;
  uint64_t *uncov = nullptr;     // [S/64] slots with an out-neighbour of another master (not covered)
  uint32_t *scol = nullptr;      // [m] push CSR columns as compute slots (shipped records; built on first use)
  uint8_t *w8 = nullptr;         // [m] push CSR weights as bytes, when every weight is below 256 (shipping)
  int ship_wbytes = 0;           // bytes per shipped weight: 0 not decided yet, 1 or 4
  uint64_t n_uncovered = 0;      // vertices not covered
  // pull row blocks (pull.hip): at most kPullNnz arcs per block, longer rows
  // split into segments merged by a fix-up pass
  grd::PullBlock *pblk = nullptr;
  uint32_t *pblk_hot = nullptr;  // per row block: first proxy of its partition's own region
  uint32_t n_pblk = 0;
  grd::PullLong *plong = nullptr;
  uint32_t n_plong = 0;
  // tiled PageRank pull (tiers.hip), built with the first pull-PageRank
  // state: the CSC regrouped by (memory partition, source tier, row), a
  // (row, tier) "pair" per row and tier it has arcs in
  bool tiers_built = false;
  int tier_mode = 0;               // 1 LDS-tiled (16-bit offsets, tiles), 2 L2-segmented (32-bit sources)
  uint32_t tier_bits = 0;          // a tier is 2^tier_bits consecutive proxies of a partition
  uint32_t tile_arcs = 0;
  uint64_t n_pairs = 0;
  uint16_t *tcol = nullptr;        // [m] mode 1: source offset within the pair's tier window
  uint32_t *tcol32 = nullptr;      // [m] mode 2: source proxy
  uint32_t *tblk_hot = nullptr;    // mode 2: first proxy of each row block's tier (its LDS window)
  uint2 *trec = nullptr;           // [pairs+1] pairs in tier order: x = first arc, y = destination
                                   // (row | 1 << 31 for a row's only pair, else its (row, tier)-order index)
  uint32_t *trptr = nullptr;       // [P+1] pairs of each proxy row, in (row, tier) order
  grd::PullBlock *tblk = nullptr;  // row blocks over pairs (r0 = first pair); nrows 0 = segment of a long pair
  uint32_t n_tblk = 0;
  grd::PullLong *tlong = nullptr;  // long pairs (row = pair index)
  uint32_t n_tlong = 0;
  grd::TierTile *ttile = nullptr;
  uint32_t n_ttile = 0;
  uint32_t *twide = nullptr;       // rows with many pairs, merged by a wave each
  uint32_t n_twide = 0;
  // mode 3, the hot / cold split: a row's arcs from its partition's kHotWin
  // hottest sources (a prefix of the row) are summed from an LDS window into
  // a per-row partial; its other arcs are pulled by the row-block kernel,
  // which merges that partial.  Hot arcs reuse tcol (16-bit window offsets),
  // tblk / tlong (their wave blocks and long rows) and ttile (each hot
  // workgroup's range of blocks, all of one partition); n_pairs = P (one
  // partial per row)
  uint32_t *hrp = nullptr;          // [P+1] hot arcs of each row
  uint32_t *corp = nullptr;         // [P+1] cold arcs of each row
  uint32_t *cocol = nullptr;        // [cold arcs] source proxies
  uint32_t *hw = nullptr;           // [hot arcs] arc weights (weighted graphs), as tcol
  uint32_t *cow = nullptr;          // [cold arcs] arc weights (weighted graphs), as cocol
  uint8_t *hw8 = nullptr;           // the same as bytes, in their place when every weight is below 256
  uint8_t *cow8 = nullptr;
  grd::PullBlock *cblk = nullptr;   // row blocks over the cold arcs
  uint32_t *cblk_hot = nullptr;     // per cold block: its LDS window base (P: none)
  uint32_t n_cblk = 0;
  grd::PullLong *clong = nullptr;
  uint32_t n_clong = 0;
  grd::DevMem tmem;
};

struct grd_state {
  grd_graph *g = nullptr;
  int algo = GRD_CC;
  grd_params prm{};
  bool initialized = false;
  uint32_t iteration = 0;
  uint64_t hi = ~0ull;  // SSSP delta-stepping bucket bound
  int last_option = 0;
  int bcast_option = -1;  // option of the last broadcast (the traversal must use the same direction)
  int xb_dense = -1, xr_dense = -1;  // last broadcast / reduce of this superstep: 0 sparse, 1 dense, -1 none
  bool sel_queued = false;  // learned: the next decision was computed on the device with the counters
  uint64_t last_frontier = 0, prev_frontier = 0, last_frontier_edges = 0, last_dirty = 0;
  uint32_t last_maxdeg = 0;
  uint64_t bytes = 0, edges = 0;
  // dominant-kernel timing (grd_params.time_kernel)
  hipEvent_t kev[2] = {nullptr, nullptr};
  int kern_pending = 0;   // 1 = pull-type launch recorded, 2 = push launch recorded
  uint64_t kern_pending_bytes = 0;
  double kern_ms = 0.0;
  uint64_t kern_launches = 0, kern_bytes = 0;
  uint32_t vbytes = 4;  // sizeof(VP) of the host instantiation: bytes counted per exchanged value
  bool wide = false;    // CC / SSSP with 8-byte properties: computed in 32 bits, widened at the boundary
  grd::DevMem mem;
  // compute side [S]
  void *prop = nullptr, *upd = nullptr;
  float *pr = nullptr;
  uint64_t *F = nullptr, *Dc = nullptr, *Pend = nullptr;  // [S/64]
  // memory side [P]
  void *mprop = nullptr, *mupd = nullptr;
  uint64_t *Du = nullptr;  // [P/64] dirty proxies
  // [P/64] CC / SSSP pull: rows whose vertex holds a value at or below the
  // superstep's smallest broadcast value (sent with a pull broadcast)
  uint64_t *Rdone = nullptr;
  bool rdone_sent = false;   // the last broadcast sent Rdone (a pull of CC / SSSP)
  bool bound_sent = false;   // it sent the smallest broadcast value (counters: fmin)
  uint32_t bound_inc = 0;    // the smallest Op::edge increment over an arc
  bool skip_pending = false; // the last pull's skipped arcs are still to be taken from the counters
  uint64_t skip_arc_bytes = 0;  // algorithmic bytes per arc of that pull
  uint8_t *bskip = nullptr;     // CC / SSSP pull: per row block, 1 = every row with arcs is done
  uint8_t *hskip = nullptr;     // the same per wave block of the hot pass (hot / cold split)
  bool hskip_on = false;        // hskip holds this pull's flags
  // wire
  uint64_t *bvB = nullptr, *bvR = nullptr;
  void *valB = nullptr, *valR = nullptr;
  // scan scratch
  uint32_t *blk = nullptr, *wrankB = nullptr, *wrankR = nullptr, *totals = nullptr;
  uint32_t *AL = nullptr;
  uint2 *heavy = nullptr;  // push: (proxy, first arc offset) of each heavy-vertex chunk
  void *partial = nullptr;  // pull: one 8-byte partial per row block (long-row segments)
  bool tiled = false;       // PR_PULL: the tiled pull runs (else the plain row-block pull)
  // tiled pull: Kahan partial per pair, then per tier-block segment; the hot /
  // cold pull keeps one 4-byte partial per row and per hot segment in them
  float2 *tpart = nullptr;
  float2 *tseg = nullptr;
  float2 *cseg = nullptr;   // hot / cold pull: partial per cold row block (long-row segments)
  double *red = nullptr;    // PR pull update: (residual, dangling) per workgroup
  // edge shipping wire (grown on demand): record lengths / offsets, the first
  // record of every tile, the records (gids) and their edge weights
  grd::DevMem ship_mem;
  uint32_t *ship_off = nullptr, *ship_tile = nullptr, *ship_buf = nullptr;
  void *ship_w = nullptr;  // the records' weights, ship_wbytes each
  uint64_t ship_kcap = 0, ship_cap = 0;
  void *ship_tmp = nullptr;  // rocPRIM scan storage (in ship_mem)
  size_t ship_tmp_bytes = 0;
  // NDPEngine inputs for the next decision: frontier size per compute
  // partition and frontier vertices not covered
  uint64_t last_fpart[grd::kMaxParts] = {};
  uint64_t last_uncovered = 0;
  grd::Counters *cnt = nullptr;   // device
  grd::Counters *hcnt = nullptr;  // pinned host
  // snapshot copies
  grd::DevMem snap_mem;
  std::vector<std::pair<void *, void *>> snap;  // (live, copy)
  std::vector<uint64_t> snap_bytes;
  uint32_t snap_iteration = 0;
  uint64_t snap_hi = 0;
  std::array<uint64_t, 8> snap_host{};  // frontier statistics, bytes and edges at the snapshot
  uint64_t snap_fpart[grd::kMaxParts] = {};
};

namespace grd {
// layout.cpp
grd_status check_row_ptr(uint32_t n, const uint32_t *rp);
grd_status build_row_blocks(grd_graph *g, const std::vector<uint32_t> &crp);
// layout_dev.hip: the same layout built by device kernels
grd_status build_graph_device(grd_ctx *ctx, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w,
                              int C, int M, const uint32_t *mirror, grd_graph **out);
grd_status build_graph(grd_ctx *ctx, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w,
                       int C, int M, const uint32_t *mirror, grd_graph **out);

void parallel_for(int nthreads, uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)> &fn);
int host_threads(const grd_ctx *ctx);
// gen.hip
grd_status generate_rmat(grd_ctx *ctx, int scale, int ef, double a, double b, double c, uint64_t seed, int sym,
                         std::vector<uint32_t> &rp, std::vector<uint32_t> &col);
grd_status generate_er(grd_ctx *ctx, uint32_t n, double deg, uint64_t seed, std::vector<uint32_t> &rp,
                       std::vector<uint32_t> &col);
void host_weights(uint32_t n, const uint32_t *rp, const uint32_t *col, uint64_t wseed, uint32_t *w, int nthreads);
// bsp.hip
grd_status state_alloc(grd_state *s);
grd_status state_init(grd_state *s);
grd_status op_broadcast(grd_state *s, int option);
grd_status op_traverse(grd_state *s, int option);
grd_status op_reduce(grd_state *s, int option);
grd_status op_update(grd_state *s);
grd_status op_select_bucket(grd_state *s, bool *any_pending);
grd_status op_aggregate_one(grd_state *s, uint32_t slot, const void *value);
grd_status op_download(grd_state *s, void *out);
grd_status op_features(grd_state *s, float *out16);
grd_status op_zero_counters(grd_state *s);
grd_status op_read_counters(grd_state *s);
// copies the per-partition frontier sizes and the uncovered count of the
// frontier just computed from the counters read back into the state
void take_frontier_parts(grd_state *s);
// ship.hip
grd_status build_ship_tables(grd_graph *g);
grd_status op_frontier_parts(grd_state *s);  // per-partition frontier counts into the counters (C > 1)
grd_status op_ship(grd_state *s);            // edge shipping: memory send + compute apply
// pull.hip: the row-block PageRank pull over any CSC cut into PullBlocks, and its
// long-row fix-up; `pre` (or null) holds a Kahan partial per row merged first
grd_status prp_pull_rows(hipStream_t st, const PullBlock *blocks, uint32_t nblk, const uint32_t *hot_base,
                         const uint32_t *rptr, const uint32_t *col, const float *mprop, uint32_t P, float *out,
                         float2 *partial, const float *pre);
grd_status prp_pull_rows_fixup(hipStream_t st, const PullLong *longs, uint32_t nlong, const float2 *partial,
                               float *out, const float *pre);
// tiers.hip
grd_status build_tiers(grd_graph *g);
// does this state pull through the hot / cold split (the BSP pull, CC / SSSP /
// reference PageRank, over a graph split by mode 3)?
inline bool bsp_hot_cold(const grd_state *s) { return s->tiled && s->g->tier_mode == 3; }
grd_status prp_pull_tiled(grd_state *s);  // pull + long-pair fix-up + merge into valR
// prpull.hip
grd_status prp_init(grd_state *s);
uint32_t prp_grid(const grd_graph *g);  // workgroups of the PR pull update
grd_status prp_iteration(grd_state *s);
grd_status prp_broadcast(grd_state *s);
grd_status prp_traverse(grd_state *s);
grd_status prp_update(grd_state *s);
// graph_io.cpp
grd_status read_graph(const std::string &path, uint32_t &n, std::vector<uint32_t> &rp, std::vector<uint32_t> &col,
                      std::vector<uint32_t> &w, bool &has_w);
grd_status read_partition_file(const std::string &path, uint32_t n, int M, std::vector<uint32_t> &mirror);
grd_status write_gr(const std::string &path, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w);
// api.cpp
int heuristic_option(const grd_state *s);
void make_features(const grd_state *s, float *f);
// mlp.hip
// device address of the selector's scratch array on `device` (current),
// taken once per device; called by grd_create
grd_status mlp_scratch_init(int device);
grd_status mlp_select(grd_ctx *ctx, const float *feat16_host, int *option, float *margin);
// Device-side selection for the next superstep from the counters of the one
// just run: computes the 16 features from the counters on the device, runs
// the MLP and queues the logits' copy to the host, so that the counters'
// readback of the superstep also brings the next decision.  prev_frontier is
// the frontier size before the superstep.
grd_status mlp_select_queue(grd_state *s, uint64_t prev_frontier);
// Reads the decision queued by mlp_select_queue from the state's counters
// (after they were read back).
void mlp_select_read(const grd_state *s, int *option, float *margin);
}  // namespace grd
