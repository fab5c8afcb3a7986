/*
 * grudon_oracle.c -- CPU oracle (test infrastructure only, see grudon_oracle.h).
 *
 * Reference files restated here (paths relative to the reference checkout):
 *   src/DistributedGraph.cpp:13-785     partition layout, lids, translation tables
 *   include/partitioning_engine/Partitioner.hpp:11-205   master/mirror rules
 *   src/GraphAlgorithm.cpp:145-701      BSP superstep and exchange
 *   include/DistributedGraph.hpp:90-214 PropertyList minUpdate/addUpdate/set/clear
 *   src/graph_algorithms/pr.cpp:20-160  PageRank operators
 *   src/graph_algorithms/cc.cpp:18-120  CC operators
 *   src/graph_algorithms/sssp.cpp:18-126 SSSP operators
 *   include/graph_algorithms/pr.hpp:6-7 DAMPING_FACTOR 0.85, TOLERANCE 1e-6
 *   include/GraphAlgorithm.hpp:10       MAX_ITERATIONS 1000 (a parameter here)
 *
 * Reference defects that are NOT reproduced: the 8x bitvector message count
 * (count = size_bytes() sent as MPI_UINT64_T, GraphAlgorithm.cpp:255,283,
 * 412,488), the thread-count division by zero and the per-thread bucket
 * overflow (generatePerThreadMatrix, GraphAlgorithm.cpp:10-66).  Exchange bytes are accounted with the correct
 * sizes: 8 * ceil(T/64) per bitvector plus sizeof(VP) per packed value.
 */
#include "grudon_oracle.h"

#include <math.h>
#include <pthread.h>
#include <time.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define U32MAX 0xFFFFFFFFu

/* ------------------------------------------------------------------ */
/* hashing + generators                                                */
/* ------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_hash(uint64_t seed, uint64_t index) { return mix64(mix64(seed) ^ index); }

static void radix_sort_u64(uint64_t *a, int64_t n)
{
  if (n < 2) return;
  uint64_t orv = 0;
  for (int64_t i = 0; i < n; i++) orv |= a[i];
  uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  uint64_t *src = a, *dst = tmp;
  for (int shift = 0; shift < 64; shift += 8) {
    if (((orv >> shift) & 0xFF) == 0) continue;
    int64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; i++) cnt[((src[i] >> shift) & 0xFF) + 1]++;
    for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFF]++] = src[i];
    uint64_t *t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(uint64_t) * (size_t)n);
  free(tmp);
}

int64_t orc_csr_from_pairs(uint32_t n, uint64_t *keys, int64_t cnt, uint32_t *row_ptr, uint32_t *col)
{
  radix_sort_u64(keys, cnt);
  memset(row_ptr, 0, sizeof(uint32_t) * ((size_t)n + 1));
  int64_t m = 0;
  for (int64_t i = 0; i < cnt; i++) {
    uint64_t k = keys[i];
    if (i > 0 && keys[i - 1] == k) continue;
    uint32_t s = (uint32_t)(k >> 32), d = (uint32_t)k;
    if (s == d) continue;
    col[m++] = d;
    row_ptr[s + 1]++;
  }
  for (uint32_t v = 0; v < n; v++) row_ptr[v + 1] += row_ptr[v];
  return m;
}

/* RMAT: level l (0 = most significant bit) of raw edge e draws 32 random bits
 * from orc_hash(seed, e*16 + l/2): the high word for even l, the low word for
 * odd l.  Quadrant thresholds are floor(p * 2^32). */
int64_t orc_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int symmetrize,
                 uint32_t *row_ptr, uint32_t *col, int64_t cap)
{
  uint32_t n = 1u << scale;
  int64_t E = (int64_t)edge_factor << scale;
  int64_t raw = symmetrize ? 2 * E : E;
  uint64_t tA = (uint64_t)(a * 4294967296.0);
  uint64_t tAB = (uint64_t)((a + b) * 4294967296.0);
  uint64_t tABC = (uint64_t)((a + b + c) * 4294967296.0);
  if (cap < raw) return -raw;
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)raw);
  for (int64_t e = 0; e < E; e++) {
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) h = orc_hash(seed, ((uint64_t)e << 4) | (uint64_t)(l >> 1));
      uint64_t r = (l & 1) ? (h & 0xFFFFFFFFull) : (h >> 32);
      uint32_t sb = (r >= tAB) ? 1u : 0u;
      uint32_t db = (r >= tA && r < tAB) || (r >= tABC) ? 1u : 0u;
      s = (s << 1) | sb;
      d = (d << 1) | db;
    }
    keys[e] = ((uint64_t)s << 32) | d;
    if (symmetrize) keys[E + e] = ((uint64_t)d << 32) | s;
  }
  int64_t m = orc_csr_from_pairs(n, keys, raw, row_ptr, col);
  free(keys);
  return m;
}

int64_t orc_er(uint32_t n, double mean_degree, uint64_t seed, uint32_t *row_ptr, uint32_t *col, int64_t cap)
{
  int64_t M = (int64_t)((double)n * mean_degree / 2.0);
  if (cap < 2 * M) return -2 * M;
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(2 * M + 1));
  for (int64_t k = 0; k < M; k++) {
    uint64_t h = orc_hash(seed, (uint64_t)k);
    uint32_t u = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
    uint32_t v = (uint32_t)(((h & 0xFFFFFFFFull) * (uint64_t)n) >> 32);
    keys[2 * k] = ((uint64_t)u << 32) | v;
    keys[2 * k + 1] = ((uint64_t)v << 32) | u;
  }
  int64_t m = orc_csr_from_pairs(n, keys, 2 * M, row_ptr, col);
  free(keys);
  return m;
}

void orc_weights(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, uint64_t wseed, uint32_t *w)
{
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
      uint32_t v = col[e];
      uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
      uint64_t h = orc_hash(wseed, ((uint64_t)lo << 32) | hi);
      w[e] = 1u + (uint32_t)(((h >> 32) * 255ull) >> 32);
    }
}

/* ------------------------------------------------------------------ */
/* tiny thread pool: one pthread per worker per parallel region        */
/* ------------------------------------------------------------------ */

typedef void (*range_fn)(void *ctx, uint64_t b, uint64_t e);
typedef struct {
  range_fn fn;
  void *ctx;
  uint64_t n, chunk;
  atomic_uint_fast64_t next;
} pf_job;

static void *pf_worker(void *arg)
{
  pf_job *j = (pf_job *)arg;
  for (;;) {
    uint64_t b = atomic_fetch_add(&j->next, j->chunk);
    if (b >= j->n) break;
    uint64_t e = b + j->chunk < j->n ? b + j->chunk : j->n;
    j->fn(j->ctx, b, e);
  }
  return NULL;
}

static void parallel_for(int nthreads, uint64_t n, uint64_t chunk, range_fn fn, void *ctx)
{
  if (n == 0) return;
  if (nthreads <= 1 || n <= chunk) {
    fn(ctx, 0, n);
    return;
  }
  pf_job j;
  j.fn = fn;
  j.ctx = ctx;
  j.n = n;
  j.chunk = chunk;
  atomic_init(&j.next, 0);
  pthread_t th[256];
  int nt = nthreads > 256 ? 256 : nthreads;
  for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, pf_worker, &j);
  pf_worker(&j);
  for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ */
/* partition layout (DistributedGraph restatement)                     */
/* ------------------------------------------------------------------ */

struct orc_part {
  uint32_t n;
  int C, M;
  uint64_t m;
  const uint32_t *grp, *gcol, *gw;
  uint32_t *mirror; /* per gid */
  uint8_t *covered; /* per gid: every out-neighbour has the same master */
  /* compute partitions: lid = rank of gid among gid % C == c (= gid / C) */
  uint32_t *nc;       /* [C] */
  uint32_t **c_gid;   /* [C][nc] */
  uint32_t **c_deg;   /* [C][nc] out-degree */
  uint32_t **c_mir;   /* [C][nc] mirror partition */
  uint32_t *T;        /* [C*M] translation table sizes, index c*M+m */
  uint32_t **c_rT;    /* [C*M][T] idx -> compute lid */
  uint32_t **c_sT;    /* [C*M][nc] compute lid -> idx or U32MAX */
  /* memory partitions: own vertices (ascending gid) then cross mirrors (ascending gid) */
  uint32_t *nm_own;   /* [M] */
  uint32_t *nm;       /* [M] */
  uint64_t *em;       /* [M] */
  uint32_t **m_gid;   /* [M][nm] */
  uint32_t **m_mas;   /* [M][nm] master partition */
  uint32_t **m_rp;    /* [M][own+1] local CSR over own vertices */
  uint32_t **m_col;   /* [M][em] memory lids */
  uint32_t **m_w;     /* [M][em] */
  uint32_t **m_rT;    /* [M*C][T] idx -> memory lid, index m*C+c */
  uint32_t **m_sT;    /* [M][nm] memory lid -> idx in pair (m, master) */
};

static void *xcalloc(size_t n, size_t s)
{
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

orc_part *orc_part_build(uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w, int C, int M,
                         const uint32_t *mirror_in)
{
  if (C < 1 || M < 1) return NULL;
  orc_part *p = (orc_part *)xcalloc(1, sizeof(orc_part));
  p->n = n;
  p->C = C;
  p->M = M;
  p->m = rp[n];
  p->grp = rp;
  p->gcol = col;
  p->gw = w;
  /* Partitioner.hpp:32-33 (naive) / :97-102, :177-178 (file); M==1 is naive (:73-85) */
  p->mirror = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  for (uint32_t g = 0; g < n; g++) {
    uint32_t mm = (M == 1 || mirror_in == NULL) ? g % (uint32_t)M : mirror_in[g];
    if (mm >= (uint32_t)M) {
      orc_part_free(p);
      return NULL;
    }
    p->mirror[g] = mm;
  }
  /* compute partitions: DistributedGraph.cpp:170-191 -> lid = gid / C */
  p->nc = (uint32_t *)xcalloc(C, sizeof(uint32_t));
  p->c_gid = (uint32_t **)xcalloc(C, sizeof(uint32_t *));
  p->c_deg = (uint32_t **)xcalloc(C, sizeof(uint32_t *));
  p->c_mir = (uint32_t **)xcalloc(C, sizeof(uint32_t *));
  for (int c = 0; c < C; c++) {
    uint32_t cnt = (n > (uint32_t)c) ? (n - 1 - (uint32_t)c) / (uint32_t)C + 1 : 0;
    p->nc[c] = cnt;
    p->c_gid[c] = (uint32_t *)xcalloc(cnt, sizeof(uint32_t));
    p->c_deg[c] = (uint32_t *)xcalloc(cnt, sizeof(uint32_t));
    p->c_mir[c] = (uint32_t *)xcalloc(cnt, sizeof(uint32_t));
    for (uint32_t l = 0; l < cnt; l++) {
      uint32_t g = l * (uint32_t)C + (uint32_t)c;
      p->c_gid[c][l] = g;
      p->c_deg[c][l] = rp[g + 1] - rp[g];
      p->c_mir[c][l] = p->mirror[g];
    }
  }
  /* cross-node mirrors: DistributedGraph.cpp:345-349, :367-371 */
  uint8_t *cross = (uint8_t *)xcalloc((size_t)M * n, 1);
  for (uint32_t u = 0; u < n; u++) {
    uint32_t mu = p->mirror[u];
    for (uint32_t e = rp[u]; e < rp[u + 1]; e++) {
      uint32_t v = col[e];
      if (p->mirror[v] != mu) cross[(size_t)mu * n + v] = 1;
    }
  }
  /* memory partitions: own lids (DistributedGraph.cpp:517-525) then cross
   * mirrors (:579-588); local CSR with dst rewritten to memory lids (:590-611) */
  p->nm_own = (uint32_t *)xcalloc(M, sizeof(uint32_t));
  p->nm = (uint32_t *)xcalloc(M, sizeof(uint32_t));
  p->em = (uint64_t *)xcalloc(M, sizeof(uint64_t));
  p->m_gid = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  p->m_mas = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  p->m_rp = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  p->m_col = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  p->m_w = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  p->m_sT = (uint32_t **)xcalloc(M, sizeof(uint32_t *));
  uint32_t *mlid = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  for (int m = 0; m < M; m++) {
    uint32_t own = 0, tot = 0;
    for (uint32_t g = 0; g < n; g++) own += (p->mirror[g] == (uint32_t)m);
    tot = own;
    for (uint32_t g = 0; g < n; g++) tot += cross[(size_t)m * n + g];
    p->nm_own[m] = own;
    p->nm[m] = tot;
    p->m_gid[m] = (uint32_t *)xcalloc(tot, sizeof(uint32_t));
    p->m_mas[m] = (uint32_t *)xcalloc(tot, sizeof(uint32_t));
    uint32_t k = 0;
    for (uint32_t g = 0; g < n; g++)
      if (p->mirror[g] == (uint32_t)m) p->m_gid[m][k++] = g;
    for (uint32_t g = 0; g < n; g++)
      if (cross[(size_t)m * n + g]) p->m_gid[m][k++] = g;
    for (uint32_t i = 0; i < n; i++) mlid[i] = U32MAX;
    for (uint32_t l = 0; l < tot; l++) {
      mlid[p->m_gid[m][l]] = l;
      p->m_mas[m][l] = p->m_gid[m][l] % (uint32_t)C;
    }
    uint64_t ecount = 0;
    for (uint32_t l = 0; l < own; l++) {
      uint32_t g = p->m_gid[m][l];
      ecount += rp[g + 1] - rp[g];
    }
    p->em[m] = ecount;
    p->m_rp[m] = (uint32_t *)xcalloc((size_t)own + 1, sizeof(uint32_t));
    p->m_col[m] = (uint32_t *)xcalloc(ecount, sizeof(uint32_t));
    p->m_w[m] = (uint32_t *)xcalloc(ecount, sizeof(uint32_t));
    uint64_t cur = 0;
    for (uint32_t l = 0; l < own; l++) {
      uint32_t g = p->m_gid[m][l];
      for (uint32_t e = rp[g]; e < rp[g + 1]; e++) {
        p->m_col[m][cur] = mlid[col[e]];
        p->m_w[m][cur] = w ? w[e] : 1u;
        cur++;
      }
      p->m_rp[m][l + 1] = (uint32_t)cur;
    }
  }
  free(mlid);
  /* translation tables */
  p->T = (uint32_t *)xcalloc((size_t)C * M, sizeof(uint32_t));
  p->c_rT = (uint32_t **)xcalloc((size_t)C * M, sizeof(uint32_t *));
  p->c_sT = (uint32_t **)xcalloc((size_t)C * M, sizeof(uint32_t *));
  p->m_rT = (uint32_t **)xcalloc((size_t)C * M, sizeof(uint32_t *));
  /* compute side (DistributedGraph.cpp:701-734): own-on-m ascending lid, then
   * cross-on-m ascending lid */
  for (int c = 0; c < C; c++)
    for (int m = 0; m < M; m++) {
      size_t pi = (size_t)c * M + m;
      uint32_t cnt = 0;
      for (uint32_t l = 0; l < p->nc[c]; l++) {
        uint32_t g = p->c_gid[c][l];
        if (p->mirror[g] == (uint32_t)m || cross[(size_t)m * n + g]) cnt++;
      }
      p->T[pi] = cnt;
      p->c_rT[pi] = (uint32_t *)xcalloc(cnt, sizeof(uint32_t));
      p->c_sT[pi] = (uint32_t *)xcalloc(p->nc[c], sizeof(uint32_t));
      for (uint32_t l = 0; l < p->nc[c]; l++) p->c_sT[pi][l] = U32MAX;
      uint32_t k = 0;
      for (uint32_t l = 0; l < p->nc[c]; l++)
        if (p->mirror[p->c_gid[c][l]] == (uint32_t)m) {
          p->c_rT[pi][k] = l;
          p->c_sT[pi][l] = k++;
        }
      for (uint32_t l = 0; l < p->nc[c]; l++) {
        uint32_t g = p->c_gid[c][l];
        if (p->mirror[g] != (uint32_t)m && cross[(size_t)m * n + g]) {
          p->c_rT[pi][k] = l;
          p->c_sT[pi][l] = k++;
        }
      }
    }
  /* memory side (DistributedGraph.cpp:644-660): vbuffer order filtered by master */
  for (int m = 0; m < M; m++) {
    p->m_sT[m] = (uint32_t *)xcalloc(p->nm[m], sizeof(uint32_t));
    uint32_t *idx = (uint32_t *)xcalloc(C, sizeof(uint32_t));
    for (int c = 0; c < C; c++) p->m_rT[(size_t)m * C + c] = (uint32_t *)xcalloc(p->T[(size_t)c * M + m], sizeof(uint32_t));
    for (uint32_t l = 0; l < p->nm[m]; l++) {
      uint32_t c = p->m_mas[m][l];
      p->m_rT[(size_t)m * C + c][idx[c]] = l;
      p->m_sT[m][l] = idx[c]++;
    }
    for (int c = 0; c < C; c++)
      if (idx[c] != p->T[(size_t)c * M + m]) {
        fprintf(stderr, "oracle: translation table size mismatch (c=%d m=%d)\n", c, m);
        abort();
      }
    free(idx);
  }
  free(cross);
  /* coverage vector (DistributedGraph.cpp:286-370, read by NDPEngine.hpp:23-26):
   * the evident intent, a vertex is covered when none of its out-arcs leaves
   * its master; the reference's guard at :338 only clears flags that are
   * already clear, so its own vector stays all true */
  p->covered = (uint8_t *)xcalloc(n, 1);
  for (uint32_t g = 0; g < n; g++) {
    uint8_t cov = 1;
    for (uint32_t e = rp[g]; e < rp[g + 1] && cov; e++)
      if (col[e] % (uint32_t)C != g % (uint32_t)C) cov = 0;
    p->covered[g] = cov;
  }
  return p;
}

void orc_part_free(orc_part *p)
{
  if (!p) return;
  for (int c = 0; c < p->C; c++) {
    if (p->c_gid) free(p->c_gid[c]);
    if (p->c_deg) free(p->c_deg[c]);
    if (p->c_mir) free(p->c_mir[c]);
  }
  for (int m = 0; m < p->M; m++) {
    if (p->m_gid) free(p->m_gid[m]);
    if (p->m_mas) free(p->m_mas[m]);
    if (p->m_rp) free(p->m_rp[m]);
    if (p->m_col) free(p->m_col[m]);
    if (p->m_w) free(p->m_w[m]);
    if (p->m_sT) free(p->m_sT[m]);
  }
  if (p->T)
    for (int i = 0; i < p->C * p->M; i++) {
      free(p->c_rT[i]);
      free(p->c_sT[i]);
      free(p->m_rT[i]);
    }
  free(p->c_gid); free(p->c_deg); free(p->c_mir); free(p->nc);
  free(p->m_gid); free(p->m_mas); free(p->m_rp); free(p->m_col); free(p->m_w); free(p->m_sT);
  free(p->nm_own); free(p->nm); free(p->em);
  free(p->T); free(p->c_rT); free(p->c_sT); free(p->m_rT);
  free(p->mirror);
  free(p->covered);
  free(p);
}

int64_t orc_part_query(const orc_part *p, int what, int i, int j, uint32_t *out)
{
  switch (what) {
  case 0: return p->T[(size_t)i * p->M + j];
  case 1: {
    size_t pi = (size_t)i * p->M + j;
    if (out)
      for (uint32_t k = 0; k < p->T[pi]; k++) out[k] = p->c_gid[i][p->c_rT[pi][k]];
    return p->T[pi];
  }
  case 2: {
    uint32_t T = p->T[(size_t)j * p->M + i];
    const uint32_t *r = p->m_rT[(size_t)i * p->C + j];
    if (out)
      for (uint32_t k = 0; k < T; k++) out[k] = p->m_gid[i][r[k]];
    return T;
  }
  case 3:
    if (out) memcpy(out, p->m_gid[i], sizeof(uint32_t) * p->nm[i]);
    return p->nm[i];
  case 4: return p->nm_own[i];
  case 5:
    if (out) memcpy(out, p->m_rp[i], sizeof(uint32_t) * ((size_t)p->nm_own[i] + 1));
    return (int64_t)p->nm_own[i] + 1;
  case 6:
    if (out) memcpy(out, p->m_col[i], sizeof(uint32_t) * p->em[i]);
    return (int64_t)p->em[i];
  case 7: return p->nc[i];
  case 8: return p->nm[i];
  case 9: return (int64_t)p->em[i];
  case 10:
    if (out)
      for (uint32_t g = 0; g < p->n; g++) out[g] = p->covered[g];
    return p->n;
  default: return -1;
  }
}

/* ------------------------------------------------------------------ */
/* atomics used by the threaded BSP (relaxed, like AtomicElement)      */
/* ------------------------------------------------------------------ */

/* returns the old value (galois::atomicMin semantics) */
static inline uint32_t amin_u32(uint32_t *a, uint32_t v)
{
  uint32_t old = __atomic_load_n(a, __ATOMIC_RELAXED);
  while (v < old && !__atomic_compare_exchange_n(a, &old, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return old;
}
static inline void aadd_f32(float *a, float v)
{
  uint32_t *ai = (uint32_t *)a;
  uint32_t old = __atomic_load_n(ai, __ATOMIC_RELAXED), nw;
  do {
    float f;
    memcpy(&f, &old, 4);
    f += v;
    memcpy(&nw, &f, 4);
  } while (!__atomic_compare_exchange_n(ai, &old, nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}
static inline void aadd_f64(double *a, double v)
{
  uint64_t *ai = (uint64_t *)a;
  uint64_t old = __atomic_load_n(ai, __ATOMIC_RELAXED), nw;
  do {
    double f;
    memcpy(&f, &old, 8);
    f += v;
    memcpy(&nw, &f, 8);
  } while (!__atomic_compare_exchange_n(ai, &old, nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}
static inline void set_flag(uint8_t *f) { __atomic_store_n(f, (uint8_t)1, __ATOMIC_RELAXED); }

/* ------------------------------------------------------------------ */
/* bitvector helpers: the exchange wire                                */
/* ------------------------------------------------------------------ */

typedef struct {
  uint64_t *bits; /* words */
  uint32_t T;
} bv_t;

/* one pair's exchange captured at one iteration (orc_bsp_pair_exchange) */
typedef struct {
  int iter, dir, c, m;
  uint32_t *gids, *vals; /* may be NULL */
  int64_t count;         /* -1 until captured */
} orc_cap;

static inline void bv_set(bv_t *b, uint32_t i) { b->bits[i >> 6] |= 1ull << (i & 63); }
static inline uint64_t bv_words(uint32_t T) { return ((uint64_t)T + 63) / 64; }
/* DynamicBitSet::getOffsets(): ascending set-bit indices */
static uint32_t bv_offsets(const bv_t *b, uint32_t *out)
{
  uint32_t k = 0;
  uint64_t W = bv_words(b->T);
  for (uint64_t w = 0; w < W; w++) {
    uint64_t x = b->bits[w];
    while (x) {
      int t = __builtin_ctzll(x);
      out[k++] = (uint32_t)(w * 64 + (uint64_t)t);
      x &= x - 1;
    }
  }
  return k;
}

/* ------------------------------------------------------------------ */
/* BSP restatement                                                     */
/* ------------------------------------------------------------------ */

#define VP uint32_t
#define SFX u32
#define IS_PR 0
#include "bsp_impl.inc"
#undef VP
#undef SFX
#undef IS_PR

#define VP float
#define VP_AADD aadd_f32
#define SFX f32
#define IS_PR 1
#include "bsp_impl.inc"
#undef VP
#undef SFX
#undef IS_PR

#undef VP_AADD
#define VP double
#define VP_AADD aadd_f64
#define SFX f64
#define IS_PR 1
#include "bsp_impl.inc"
#undef VP
#undef SFX
#undef IS_PR

int orc_bsp_u32(orc_part *p, int algo, uint32_t source, int weighted, uint32_t delta, int max_iter, int nthreads,
                uint32_t *out, uint64_t *stats, uint32_t *trace)
{
  if (algo != ORC_CC && algo != ORC_SSSP) return -1;
  if (algo == ORC_SSSP && source >= p->n) return -2;
  double *dummy = NULL;
  return bsp_run_u32(p, algo, source, weighted, delta, max_iter, nthreads, 0, out, dummy, stats, trace, NULL, NULL);
}

int orc_bsp_pr(orc_part *p, int fp64, int max_iter, int nthreads, double *out, uint64_t *stats, uint32_t *trace)
{
  return orc_bsp_pr_sched(p, fp64, max_iter, nthreads, 0, out, stats, trace, NULL);
}

int orc_bsp_u32_sched(orc_part *p, int algo, uint32_t source, int weighted, uint32_t delta, int max_iter, int nthreads,
                      int sched, uint32_t *out, uint64_t *stats, uint32_t *trace, uint8_t *votes)
{
  if (algo != ORC_CC && algo != ORC_SSSP) return -1;
  if (algo == ORC_SSSP && source >= p->n) return -2;
  if (sched < 0 || sched > 2) return -3;
  return bsp_run_u32(p, algo, source, weighted, delta, max_iter, nthreads, sched, out, NULL, stats, trace, votes,
                     NULL);
}

static int cmp_pair_u64(const void *a, const void *b)
{
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

int64_t orc_bsp_pair_exchange(orc_part *p, int algo, uint32_t source, int weighted, uint32_t delta, int iter, int dir,
                              int c, int m, uint32_t *gids, uint32_t *vals)
{
  if (algo != ORC_CC && algo != ORC_SSSP) return -1;
  if (algo == ORC_SSSP && source >= p->n) return -2;
  if (c < 0 || m < 0 || c >= p->C || m >= p->M || (dir != 0 && dir != 1) || iter < 0) return -3;
  orc_cap cap = {iter, dir, c, m, NULL, NULL, -1};
  uint64_t T = p->T[(size_t)c * p->M + m];
  uint64_t *kv = (uint64_t *)xcalloc(T ? T : 1, sizeof(uint64_t));
  cap.gids = (uint32_t *)xcalloc(T ? T : 1, sizeof(uint32_t));
  cap.vals = (uint32_t *)xcalloc(T ? T : 1, sizeof(uint32_t));
  uint32_t *out = (uint32_t *)xcalloc(p->n, sizeof(uint32_t));
  uint64_t stats[4];
  bsp_run_u32(p, algo, source, weighted, delta, iter + 1, 1, 0, out, NULL, stats, NULL, NULL, &cap);
  int64_t k = cap.count < 0 ? 0 : cap.count;
  if (k > 0) {
    /* ascending gid, values alongside (gid << 32 | value) */
    for (int64_t i = 0; i < k; i++) kv[i] = ((uint64_t)cap.gids[i] << 32) | cap.vals[i];
    qsort(kv, (size_t)k, sizeof(uint64_t), cmp_pair_u64);
    for (int64_t i = 0; i < k; i++) {
      if (gids) gids[i] = (uint32_t)(kv[i] >> 32);
      if (vals) vals[i] = (uint32_t)kv[i];
    }
  }
  free(kv);
  free(cap.gids);
  free(cap.vals);
  free(out);
  return k;
}

int orc_bsp_pr_sched(orc_part *p, int fp64, int max_iter, int nthreads, int sched, double *out, uint64_t *stats,
                     uint32_t *trace, uint8_t *votes)
{
  if (sched < 0 || sched > 2) return -3;
  if (fp64) return bsp_run_f64(p, ORC_PR, 0, 0, 0, max_iter, nthreads, sched, NULL, out, stats, trace, votes, NULL);
  return bsp_run_f32(p, ORC_PR, 0, 0, 0, max_iter, nthreads, sched, NULL, out, stats, trace, votes, NULL);
}

/* ------------------------------------------------------------------ */
/* closed forms                                                        */
/* ------------------------------------------------------------------ */

void orc_bfs(uint32_t n, const uint32_t *rp, const uint32_t *col, uint32_t src, uint32_t *dist)
{
  for (uint32_t i = 0; i < n; i++) dist[i] = U32MAX;
  uint32_t *q = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  uint32_t h = 0, t = 0;
  dist[src] = 0;
  q[t++] = src;
  while (h < t) {
    uint32_t u = q[h++];
    for (uint32_t e = rp[u]; e < rp[u + 1]; e++) {
      uint32_t v = col[e];
      if (dist[v] == U32MAX) {
        dist[v] = dist[u] + 1;
        q[t++] = v;
      }
    }
  }
  free(q);
}

/* binary heap Dijkstra with lazy deletion */
typedef struct {
  uint64_t key; /* dist << 32 | vertex */
} hitem;
static void hpush(hitem *h, uint64_t *sz, uint64_t key)
{
  uint64_t i = (*sz)++;
  h[i].key = key;
  while (i > 0) {
    uint64_t pa = (i - 1) / 2;
    if (h[pa].key <= h[i].key) break;
    hitem t = h[pa]; h[pa] = h[i]; h[i] = t;
    i = pa;
  }
}
static uint64_t hpop(hitem *h, uint64_t *sz)
{
  uint64_t top = h[0].key;
  h[0] = h[--(*sz)];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, s = i;
    if (l < *sz && h[l].key < h[s].key) s = l;
    if (r < *sz && h[r].key < h[s].key) s = r;
    if (s == i) break;
    hitem t = h[s]; h[s] = h[i]; h[i] = t;
    i = s;
  }
  return top;
}

void orc_dijkstra(uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w, uint32_t src, uint32_t *dist)
{
  for (uint32_t i = 0; i < n; i++) dist[i] = U32MAX;
  uint64_t cap = (uint64_t)rp[n] + 1;
  hitem *h = (hitem *)xcalloc(cap, sizeof(hitem));
  uint64_t sz = 0;
  dist[src] = 0;
  hpush(h, &sz, ((uint64_t)0 << 32) | src);
  while (sz) {
    uint64_t k = hpop(h, &sz);
    uint32_t d = (uint32_t)(k >> 32), u = (uint32_t)k;
    if (d != dist[u]) continue;
    for (uint32_t e = rp[u]; e < rp[u + 1]; e++) {
      uint64_t nd = (uint64_t)d + (w ? w[e] : 1u);
      if (nd >= U32MAX) nd = U32MAX;
      uint32_t v = col[e];
      if (nd < dist[v]) {
        dist[v] = (uint32_t)nd;
        hpush(h, &sz, (nd << 32) | v);
      }
    }
  }
  free(h);
}

static uint32_t uf_find(uint32_t *pa, uint32_t x)
{
  while (pa[x] != x) {
    pa[x] = pa[pa[x]];
    x = pa[x];
  }
  return x;
}

void orc_cc_unionfind(uint32_t n, const uint32_t *rp, const uint32_t *col, uint32_t *label)
{
  uint32_t *pa = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; i++) pa[i] = i;
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t e = rp[u]; e < rp[u + 1]; e++) {
      uint32_t a = uf_find(pa, u), b = uf_find(pa, col[e]);
      if (a < b) pa[b] = a;
      else if (b < a) pa[a] = b;
    }
  for (uint32_t i = 0; i < n; i++) label[i] = uf_find(pa, i); /* root = min gid of the set */
  free(pa);
}

int orc_pr_pull(uint32_t n, const uint32_t *rp, const uint32_t *col, double d, int iters, double tol, double *r,
                double *last_residual)
{
  double *nx = (double *)xcalloc(n, sizeof(double));
  for (uint32_t i = 0; i < n; i++) r[i] = 1.0 / (double)n;
  int it = 0, cap = iters > 0 ? iters : 10000;
  double res = 0.0;
  while (it < cap) {
    double D = 0.0;
    for (uint32_t u = 0; u < n; u++)
      if (rp[u + 1] == rp[u]) D += r[u];
    for (uint32_t v = 0; v < n; v++) nx[v] = 0.0;
    for (uint32_t u = 0; u < n; u++) {
      uint32_t deg = rp[u + 1] - rp[u];
      if (!deg) continue;
      double cu = r[u] / (double)deg;
      for (uint32_t e = rp[u]; e < rp[u + 1]; e++) nx[col[e]] += cu;
    }
    res = 0.0;
    double base = (1.0 - d) / (double)n + d * D / (double)n;
    for (uint32_t v = 0; v < n; v++) {
      double x = base + d * nx[v];
      res += fabs(x - r[v]);
      r[v] = x;
    }
    it++;
    if (iters <= 0 && res < tol) break;
  }
  if (last_residual) *last_residual = res;
  free(nx);
  return it;
}

/* ---- multithreaded FP64 pull PageRank (same recurrence as orc_pr_pull) ---- */

typedef struct {
  uint32_t n;
  const uint32_t *trp, *tcol, *deg;
  const double *r;
  double *contrib, *nx;
  double base, d;
} prmt_ctx;

static void prmt_contrib(void *vc, uint64_t b, uint64_t e)
{
  prmt_ctx *c = (prmt_ctx *)vc;
  for (uint64_t u = b; u < e; u++) c->contrib[u] = c->deg[u] ? c->r[u] / (double)c->deg[u] : 0.0;
}

static void prmt_pull(void *vc, uint64_t b, uint64_t e)
{
  prmt_ctx *c = (prmt_ctx *)vc;
  for (uint64_t v = b; v < e; v++) {
    double s = 0.0;
    for (uint32_t x = c->trp[v]; x < c->trp[v + 1]; x++) s += c->contrib[c->tcol[x]];
    c->nx[v] = c->base + c->d * s;
  }
}

int orc_pr_pull_mt(uint32_t n, const uint32_t *rp, const uint32_t *col, double d, int iters, double tol, double *r,
                   double *last_residual, int nthreads)
{
  uint64_t m = rp[n];
  uint32_t *trp = (uint32_t *)xcalloc((size_t)n + 1, sizeof(uint32_t));
  uint32_t *tcol = (uint32_t *)xcalloc(m, sizeof(uint32_t));
  uint32_t *deg = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  for (uint64_t x = 0; x < m; x++) trp[col[x] + 1]++;
  for (uint32_t v = 0; v < n; v++) trp[v + 1] += trp[v];
  uint32_t *cur = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  memcpy(cur, trp, sizeof(uint32_t) * n);
  for (uint32_t u = 0; u < n; u++) {
    deg[u] = rp[u + 1] - rp[u];
    for (uint32_t x = rp[u]; x < rp[u + 1]; x++) tcol[cur[col[x]]++] = u;
  }
  free(cur);
  double *contrib = (double *)xcalloc(n, sizeof(double));
  double *nx = (double *)xcalloc(n, sizeof(double));
  for (uint32_t i = 0; i < n; i++) r[i] = 1.0 / (double)n;
  prmt_ctx c = {n, trp, tcol, deg, r, contrib, nx, 0.0, d};
  int it = 0, cap = iters > 0 ? iters : 10000;
  double res = 0.0;
  while (it < cap) {
    double D = 0.0;
    for (uint32_t u = 0; u < n; u++)
      if (!deg[u]) D += r[u];
    c.base = (1.0 - d) / (double)n + d * D / (double)n;
    parallel_for(nthreads, n, 65536, prmt_contrib, &c);
    parallel_for(nthreads, n, 16384, prmt_pull, &c);
    res = 0.0;
    for (uint32_t v = 0; v < n; v++) {
      res += fabs(nx[v] - r[v]);
      r[v] = nx[v];
    }
    it++;
    if (iters <= 0 && res < tol) break;
  }
  if (last_residual) *last_residual = res;
  free(trp);
  free(tcol);
  free(deg);
  free(contrib);
  free(nx);
  return it;
}

/* ------------------------------------------------------------------ */
/* full-size checks and the timed CPU pull (bench.py)                  */
/* ------------------------------------------------------------------ */

struct orc_pull_plan {
  uint32_t n;
  int nthreads;
  uint32_t *trp, *tcol, *deg;
  double *contrib, *nx;
};

typedef struct {
  const uint32_t *rp, *col;
  uint32_t *trp, *deg;
  _Atomic uint32_t *cur;
  uint32_t *tcol;
} tr_ctx;

static void tr_count(void *vc, uint64_t b, uint64_t e)
{
  tr_ctx *c = (tr_ctx *)vc;
  for (uint64_t u = b; u < e; u++) {
    c->deg[u] = c->rp[u + 1] - c->rp[u];
    for (uint32_t x = c->rp[u]; x < c->rp[u + 1]; x++) atomic_fetch_add_explicit(&c->cur[c->col[x]], 1u, memory_order_relaxed);
  }
}

static void tr_fill(void *vc, uint64_t b, uint64_t e)
{
  tr_ctx *c = (tr_ctx *)vc;
  for (uint64_t u = b; u < e; u++)
    for (uint32_t x = c->rp[u]; x < c->rp[u + 1]; x++)
      c->tcol[atomic_fetch_add_explicit(&c->cur[c->col[x]], 1u, memory_order_relaxed)] = (uint32_t)u;
}

static int cmp_u32(const void *a, const void *b)
{
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

static void tr_sort(void *vc, uint64_t b, uint64_t e)
{
  tr_ctx *c = (tr_ctx *)vc;
  for (uint64_t v = b; v < e; v++) {
    uint32_t lo = c->trp[v], hi = c->trp[v + 1];
    if (hi - lo > 1) qsort(c->tcol + lo, hi - lo, sizeof(uint32_t), cmp_u32);
  }
}

orc_pull_plan *orc_pull_plan_build(uint32_t n, const uint32_t *rp, const uint32_t *col, int nthreads)
{
  orc_pull_plan *p = (orc_pull_plan *)xcalloc(1, sizeof(orc_pull_plan));
  const uint64_t m = rp[n];
  p->n = n;
  p->nthreads = nthreads;
  p->trp = (uint32_t *)xcalloc((size_t)n + 1, sizeof(uint32_t));
  p->tcol = (uint32_t *)xcalloc(m, sizeof(uint32_t));
  p->deg = (uint32_t *)xcalloc(n, sizeof(uint32_t));
  p->contrib = (double *)xcalloc(n, sizeof(double));
  p->nx = (double *)xcalloc(n, sizeof(double));
  _Atomic uint32_t *cur = (_Atomic uint32_t *)xcalloc(n, sizeof(uint32_t));
  tr_ctx c = {rp, col, p->trp, p->deg, cur, p->tcol};
  parallel_for(nthreads, n, 16384, tr_count, &c);
  /* in-degree counts -> row starts; the cursors restart at the row starts */
  uint32_t acc = 0;
  for (uint32_t v = 0; v < n; v++) {
    uint32_t k = atomic_load_explicit(&cur[v], memory_order_relaxed);
    p->trp[v] = acc;
    atomic_store_explicit(&cur[v], acc, memory_order_relaxed);
    acc += k;
  }
  p->trp[n] = acc;
  parallel_for(nthreads, n, 16384, tr_fill, &c);
  /* ascending sources per row: the summation order of orc_pr_pull_mt */
  parallel_for(nthreads, n, 4096, tr_sort, &c);
  free((void *)cur);
  return p;
}

int orc_pull_plan_run(orc_pull_plan *p, double d, int iters, double tol, double *r, double *last_residual,
                      double *seconds)
{
  struct timespec t0, t1;
  const uint32_t n = p->n;
  for (uint32_t i = 0; i < n; i++) r[i] = 1.0 / (double)n;
  prmt_ctx c = {n, p->trp, p->tcol, p->deg, r, p->contrib, p->nx, 0.0, d};
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int it = 0, cap = iters > 0 ? iters : 10000;
  double res = 0.0;
  while (it < cap) {
    double D = 0.0;
    for (uint32_t u = 0; u < n; u++)
      if (!p->deg[u]) D += r[u];
    c.base = (1.0 - d) / (double)n + d * D / (double)n;
    parallel_for(p->nthreads, n, 65536, prmt_contrib, &c);
    parallel_for(p->nthreads, n, 16384, prmt_pull, &c);
    res = 0.0;
    for (uint32_t v = 0; v < n; v++) {
      res += fabs(p->nx[v] - r[v]);
      r[v] = p->nx[v];
    }
    it++;
    if (iters <= 0 && res < tol) break;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (last_residual) *last_residual = res;
  if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  return it;
}

uint64_t orc_pull_plan_arcs(const orc_pull_plan *p) { return p->trp[p->n]; }

void orc_pull_plan_free(orc_pull_plan *p)
{
  if (!p) return;
  free(p->trp);
  free(p->tcol);
  free(p->deg);
  free(p->contrib);
  free(p->nx);
  free(p);
}

/* parallel union-find: a root is linked only under a smaller root, so the
 * root of every set is its smallest gid */
typedef struct {
  const uint32_t *rp, *col;
  _Atomic uint32_t *pa;
  uint32_t *label;
} uf_ctx;

static uint32_t ufm_find(_Atomic uint32_t *pa, uint32_t x)
{
  for (;;) {
    uint32_t p = atomic_load_explicit(&pa[x], memory_order_relaxed);
    if (p == x) return x;
    uint32_t gp = atomic_load_explicit(&pa[p], memory_order_relaxed);
    if (gp != p) atomic_compare_exchange_weak_explicit(&pa[x], &p, gp, memory_order_relaxed, memory_order_relaxed);
    x = gp;
  }
}

static void uf_union_range(void *vc, uint64_t b, uint64_t e)
{
  uf_ctx *c = (uf_ctx *)vc;
  for (uint64_t u = b; u < e; u++)
    for (uint32_t x = c->rp[u]; x < c->rp[u + 1]; x++) {
      uint32_t a = (uint32_t)u, v = c->col[x];
      for (;;) {
        uint32_t ra = ufm_find(c->pa, a), rb = ufm_find(c->pa, v);
        if (ra == rb) break;
        uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
        uint32_t expect = hi;
        if (atomic_compare_exchange_strong_explicit(&c->pa[hi], &expect, lo, memory_order_relaxed,
                                                    memory_order_relaxed))
          break;
        a = hi;
        v = lo;
      }
    }
}

static void uf_label_range(void *vc, uint64_t b, uint64_t e)
{
  uf_ctx *c = (uf_ctx *)vc;
  for (uint64_t v = b; v < e; v++) c->label[v] = ufm_find(c->pa, (uint32_t)v);
}

void orc_cc_unionfind_mt(uint32_t n, const uint32_t *rp, const uint32_t *col, uint32_t *label, int nthreads)
{
  _Atomic uint32_t *pa = (_Atomic uint32_t *)xcalloc(n, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; i++) atomic_init(&pa[i], i);
  uf_ctx c = {rp, col, pa, label};
  parallel_for(nthreads, n, 4096, uf_union_range, &c);
  parallel_for(nthreads, n, 65536, uf_label_range, &c);
  free((void *)pa);
}

/* Bellman equations of a distance vector (positive weights, or +1 per arc
 * when w is NULL): dist[src] = 0; no arc u -> v offers dist[u] + w < dist[v]
 * (an unreached 2^32 - 1 below a reached vertex counts); every reached
 * v != src has an arc u -> v with dist[u] + w == dist[v].  These hold for
 * the shortest-path distances and for no other vector. */
typedef struct {
  const uint32_t *rp, *col, *w, *dist;
  uint8_t *tight;
  _Atomic uint64_t bad;
} bc_ctx;

static void bc_arcs(void *vc, uint64_t b, uint64_t e)
{
  bc_ctx *c = (bc_ctx *)vc;
  uint64_t bad = 0;
  for (uint64_t u = b; u < e; u++) {
    const uint32_t du = c->dist[u];
    if (du == 0xFFFFFFFFu) continue;
    for (uint32_t x = c->rp[u]; x < c->rp[u + 1]; x++) {
      const uint32_t v = c->col[x];
      const uint64_t cand = (uint64_t)du + (c->w ? c->w[x] : 1u);
      if ((uint64_t)c->dist[v] > cand) bad++;
      else if ((uint64_t)c->dist[v] == cand) c->tight[v] = 1;
    }
  }
  if (bad) atomic_fetch_add(&c->bad, bad);
}

uint64_t orc_sssp_check_mt(uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w,
                           const uint32_t *dist, uint32_t src, int nthreads)
{
  uint8_t *tight = (uint8_t *)xcalloc(n, 1);
  bc_ctx c = {rp, col, w, dist, tight, 0};
  atomic_init(&c.bad, 0);
  parallel_for(nthreads, n, 4096, bc_arcs, &c);
  uint64_t bad = atomic_load(&c.bad);
  if (src >= n || dist[src] != 0) bad++;
  for (uint32_t v = 0; v < n; v++)
    if (v != src && dist[v] != 0xFFFFFFFFu && !tight[v]) bad++;
  free(tight);
  return bad;
}
