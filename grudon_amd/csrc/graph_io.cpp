// graph_io.cpp -- graph and partition-file readers.
//
// Galois binary .gr (the reference's only input, DistributedGraph.cpp:49,
// :571 via galois::graphs::readGraph), version 1 (u32 destinations) and
// version 2 (u64 destinations): four u64 header words (version, edge-data
// size, nodes, edges), u64 row-end offsets, destinations (v1 padded to 8
// bytes), then the edge data.  MatrixMarket coordinate .mtx (the input that
// scripts/convert_and_partition.sh converts) is read directly, 1-based,
// "symmetric" adds the mirrored arc.  Partition files hold one memory
// partition id per line (Partitioner.hpp:89-105).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "grd_internal.hpp"

namespace grd {

// CSR from (src << 32 | dst) keys.  File inputs keep duplicates and
// self-loops, as the reference traverses whatever the file holds; rows are
// ordered by a stable sort on the source only.
grd_status csr_from_pairs(uint32_t n, std::vector<uint64_t> &keys, std::vector<uint32_t> *wts,
                          std::vector<uint32_t> &rp, std::vector<uint32_t> &col, std::vector<uint32_t> &w)
{
  std::vector<uint64_t> idx(keys.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return (keys[a] >> 32) < (keys[b] >> 32); });
  rp.assign((size_t)n + 1, 0);
  col.clear();
  w.clear();
  for (uint64_t i : idx) {
    uint64_t k = keys[i];
    uint32_t s = (uint32_t)(k >> 32), d = (uint32_t)k;
    col.push_back(d);
    if (wts) w.push_back((*wts)[i]);
    rp[s + 1]++;
  }
  for (uint32_t v = 0; v < n; v++) rp[v + 1] += rp[v];
  if (col.size() >= 0xFFFFFFF0ull) return fail(GRD_ERR_ARG, "graph: too many arcs for u32 offsets");
  return GRD_OK;
}

static grd_status read_gr(const std::string &path, uint32_t &n, std::vector<uint32_t> &rp, std::vector<uint32_t> &col,
                          std::vector<uint32_t> &w, bool &has_w)
{
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return fail(GRD_ERR_IO, "cannot open graph file " + path);
  uint64_t hdr[4];
  if (fread(hdr, 8, 4, f) != 4) {
    fclose(f);
    return fail(GRD_ERR_IO, "truncated .gr header in " + path);
  }
  const uint64_t ver = hdr[0], esz = hdr[1], nn = hdr[2], ne = hdr[3];
  if ((ver != 1 && ver != 2) || nn >= 0xFFFFFFFFull || ne >= 0xFFFFFFF0ull) {
    fclose(f);
    return fail(GRD_ERR_IO, "unsupported .gr file (version/size) " + path);
  }
  // edge data: none, or the u32 weights of the reference's uint32 edge type
  // (graph-convert --edgeType=uint32, scripts/convert_and_partition.sh:22-29)
  if (esz != 0 && esz != 4) {
    fclose(f);
    return fail(GRD_ERR_IO, "unsupported .gr edge data of " + std::to_string(esz) + " bytes (0 or 4) in " + path);
  }
  n = (uint32_t)nn;
  std::vector<uint64_t> ends(nn);
  if (nn && fread(ends.data(), 8, nn, f) != nn) {
    fclose(f);
    return fail(GRD_ERR_IO, "truncated .gr offsets in " + path);
  }
  std::vector<uint64_t> keys(ne);
  std::vector<uint32_t> wts;
  std::vector<uint32_t> dst32;
  if (ver == 1) {
    dst32.resize(ne);
    if (ne && fread(dst32.data(), 4, ne, f) != ne) {
      fclose(f);
      return fail(GRD_ERR_IO, "truncated .gr edges in " + path);
    }
    if (ne & 1) {
      uint32_t pad;
      if (fread(&pad, 4, 1, f) != 1 && esz) {
        fclose(f);
        return fail(GRD_ERR_IO, "truncated .gr padding in " + path);
      }
    }
  } else {
    std::vector<uint64_t> d64(ne);
    if (ne && fread(d64.data(), 8, ne, f) != ne) {
      fclose(f);
      return fail(GRD_ERR_IO, "truncated .gr edges in " + path);
    }
    dst32.resize(ne);
    for (uint64_t i = 0; i < ne; i++) {
      // range-check before narrowing: 2^32 + v must not alias vertex v
      if (d64[i] >= nn) {
        fclose(f);
        return fail(GRD_ERR_IO, "corrupt .gr destination in " + path);
      }
      dst32[i] = (uint32_t)d64[i];
    }
  }
  has_w = (esz == 4);
  if (has_w) {
    wts.resize(ne);
    if (ne && fread(wts.data(), 4, ne, f) != ne) {
      fclose(f);
      return fail(GRD_ERR_IO, "truncated .gr edge data in " + path);
    }
  }
  fclose(f);
  uint64_t e = 0;
  for (uint64_t v = 0; v < nn; v++) {
    if (ends[v] < e || ends[v] > ne) return fail(GRD_ERR_IO, "corrupt .gr offsets in " + path);
    for (; e < ends[v]; e++) {
      if (dst32[e] >= nn) return fail(GRD_ERR_IO, "corrupt .gr destination in " + path);
      keys[e] = ((uint64_t)v << 32) | dst32[e];
    }
  }
  // the last node's out-index must close the arc array
  if (e != ne) return fail(GRD_ERR_IO, "corrupt .gr offsets (the last row ends before the last arc) in " + path);
  return csr_from_pairs(n, keys, has_w ? &wts : nullptr, rp, col, w);
}

static grd_status read_mtx(const std::string &path, uint32_t &n, std::vector<uint32_t> &rp, std::vector<uint32_t> &col,
                           std::vector<uint32_t> &w, bool &has_w)
{
  std::ifstream in(path);
  if (!in) return fail(GRD_ERR_IO, "cannot open graph file " + path);
  std::string line;
  bool symmetric = false, pattern = true, header = false;
  uint64_t rows = 0, cols = 0, nnz = 0;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    if (line[0] == '%') {
      if (line.rfind("%%MatrixMarket", 0) == 0) {
        std::string l = line;
        std::transform(l.begin(), l.end(), l.begin(), ::tolower);
        symmetric = l.find("symmetric") != std::string::npos;
        pattern = l.find("pattern") != std::string::npos;
      }
      continue;
    }
    std::istringstream ss(line);
    if (!(ss >> rows >> cols >> nnz)) return fail(GRD_ERR_IO, "bad .mtx size line in " + path);
    header = true;
    break;
  }
  if (!header || rows != cols || rows >= 0xFFFFFFFFull) return fail(GRD_ERR_IO, "bad or non-square .mtx " + path);
  n = (uint32_t)rows;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> wts;
  keys.reserve(symmetric ? 2 * nnz : nnz);
  has_w = !pattern;
  for (uint64_t k = 0; k < nnz; k++) {
    if (!std::getline(in, line)) return fail(GRD_ERR_IO, "truncated .mtx " + path);
    if (line.empty() || line[0] == '%') {
      k--;
      continue;
    }
    std::istringstream ss(line);
    uint64_t i, j;
    double val = 1.0;
    if (!(ss >> i >> j)) return fail(GRD_ERR_IO, "bad .mtx entry in " + path);
    if (!pattern) ss >> val;
    if (i < 1 || j < 1 || i > rows || j > cols) return fail(GRD_ERR_IO, "out-of-range .mtx entry in " + path);
    uint32_t wv = (uint32_t)std::max(1.0, std::min(4294967295.0, val));
    keys.push_back(((i - 1) << 32) | (j - 1));
    if (has_w) wts.push_back(wv);
    if (symmetric && i != j) {
      keys.push_back(((j - 1) << 32) | (i - 1));
      if (has_w) wts.push_back(wv);
    }
  }
  return csr_from_pairs(n, keys, has_w ? &wts : nullptr, rp, col, w);
}

grd_status read_graph(const std::string &path, uint32_t &n, std::vector<uint32_t> &rp, std::vector<uint32_t> &col,
                      std::vector<uint32_t> &w, bool &has_w)
{
  auto ends_with = [&](const char *suf) {
    size_t l = strlen(suf);
    return path.size() >= l && path.compare(path.size() - l, l, suf) == 0;
  };
  if (ends_with(".mtx")) return read_mtx(path, n, rp, col, w, has_w);
  return read_gr(path, n, rp, col, w, has_w);
}

grd_status read_partition_file(const std::string &path, uint32_t n, int M, std::vector<uint32_t> &mirror)
{
  std::ifstream in(path);
  if (!in) return fail(GRD_ERR_IO, "Failed to open partitioning scheme file: " + path);
  mirror.assign(n, 0);
  for (uint32_t i = 0; i < n; i++) {
    long long x;
    if (!(in >> x)) return fail(GRD_ERR_IO, "partition file " + path + " has fewer than n entries");
    if (x < 0 || x >= M) return fail(GRD_ERR_IO, "partition file " + path + ": id out of range [0, M)");
    mirror[i] = (uint32_t)x;
  }
  return GRD_OK;
}

grd_status write_gr(const std::string &path, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w)
{
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return fail(GRD_ERR_IO, "cannot write " + path);
  uint64_t ne = rp[n];
  uint64_t hdr[4] = {1, w ? 4ull : 0ull, n, ne};
  fwrite(hdr, 8, 4, f);
  std::vector<uint64_t> ends(n);
  for (uint32_t v = 0; v < n; v++) ends[v] = rp[v + 1];
  fwrite(ends.data(), 8, n, f);
  fwrite(col, 4, ne, f);
  if (ne & 1) {
    uint32_t pad = 0;
    fwrite(&pad, 4, 1, f);
  }
  if (w) fwrite(w, 4, ne, f);
  bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? GRD_OK : fail(GRD_ERR_IO, "write error on " + path);
}

}  // namespace grd
