// DistributedGraph.cpp -- graph loading and the resident partition layout.
#include "DistributedGraph.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace {

std::vector<std::string> split(const std::string &s, char sep)
{
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, sep)) out.push_back(item);
  return out;
}

void check(grd_status st, const std::string &what)
{
  if (st != GRD_OK) throw std::runtime_error(what + ": " + grd_last_error());
}

}  // namespace

HostGraph load_host_graph(Exchange &net, const std::string &spec)
{
  HostGraph g;
  const auto parts = split(spec, ':');
  if (!parts.empty() && parts[0] == "rmat" && parts.size() >= 2) {
    const int scale = std::atoi(parts[1].c_str());
    const int ef = parts.size() > 2 ? std::atoi(parts[2].c_str()) : 16;
    const uint64_t seed = parts.size() > 3 ? std::strtoull(parts[3].c_str(), nullptr, 10) : 1;
    const bool sym = parts.size() > 4 && (parts[4] == "sym" || parts[4] == "1");
    const uint64_t wseed = parts.size() > 5 ? std::strtoull(parts[5].c_str(), nullptr, 10) : 0;
    if (scale < 1 || scale > 30) throw std::runtime_error("rmat spec: scale out of range");
    g.n = 1u << scale;
    const uint64_t cap = ((uint64_t)ef << scale) * (sym ? 2 : 1);
    g.row_ptr.resize((size_t)g.n + 1);
    g.col.resize(cap);
    uint64_t m = 0;
    check(grd_generate_rmat(net.context(), scale, ef, 0.57, 0.19, 0.19, seed, sym ? 1 : 0, g.row_ptr.data(),
                            g.col.data(), cap, &m),
          "rmat generator");
    g.col.resize(m);
    if (wseed) {
      g.weights.resize(m);
      check(grd_edge_weights(net.context(), g.n, g.row_ptr.data(), g.col.data(), wseed, g.weights.data()),
            "edge weights");
    }
    return g;
  }
  if (!parts.empty() && parts[0] == "er" && parts.size() >= 3) {
    const uint32_t n = (uint32_t)std::strtoul(parts[1].c_str(), nullptr, 10);
    const double deg = std::atof(parts[2].c_str());
    const uint64_t seed = parts.size() > 3 ? std::strtoull(parts[3].c_str(), nullptr, 10) : 1;
    g.n = n;
    const uint64_t cap = (uint64_t)((double)n * deg / 2.0) * 2 + 2;
    g.row_ptr.resize((size_t)n + 1);
    g.col.resize(cap);
    uint64_t m = 0;
    check(grd_generate_er(net.context(), n, deg, seed, g.row_ptr.data(), g.col.data(), cap, &m), "er generator");
    g.col.resize(m);
    const uint64_t wseed = parts.size() > 4 ? std::strtoull(parts[4].c_str(), nullptr, 10) : 0;
    if (wseed) {
      g.weights.resize(m);
      check(grd_edge_weights(net.context(), g.n, g.row_ptr.data(), g.col.data(), wseed, g.weights.data()),
            "edge weights");
    }
    return g;
  }
  // a file: .gr (Galois binary) or .mtx
  uint32_t n = 0;
  uint64_t m = 0;
  int hw = 0;
  check(grd_graph_read(spec.c_str(), &n, &m, nullptr, nullptr, nullptr, &hw), "reading " + spec);
  g.n = n;
  g.row_ptr.resize((size_t)n + 1);
  g.col.resize(m);
  if (hw) g.weights.resize(m);
  check(grd_graph_read(spec.c_str(), &n, &m, g.row_ptr.data(), g.col.data(), hw ? g.weights.data() : nullptr, &hw),
        "reading " + spec);
  return g;
}

std::vector<uint32_t> load_partition_file(const std::string &path, uint32_t n, size_t num_memory)
{
  std::ifstream in(path);
  if (!in) throw std::runtime_error("Failed to open partitioning scheme file: " + path);
  std::vector<uint32_t> mirror(n);
  for (uint32_t i = 0; i < n; i++) {
    long long x;
    if (!(in >> x)) throw std::runtime_error("partition file " + path + " has fewer than n entries");
    if (x < 0 || (size_t)x >= num_memory) throw std::runtime_error("partition file " + path + ": id out of range");
    mirror[i] = (uint32_t)x;
  }
  return mirror;
}

DistributedGraph::DistributedGraph(Exchange &net, const HostGraph &g, size_t num_compute, size_t num_memory,
                                   const std::string &partitioning_scheme_file)
    : host_(g), num_compute_(num_compute), num_memory_(num_memory)
{
  // Partitioner.hpp:73-85: with one memory partition the file is ignored
  if (!partitioning_scheme_file.empty() && num_memory > 1)
    mirror_ = load_partition_file(partitioning_scheme_file, g.n, num_memory);
  check(grd_graph_from_csr(net.context(), g.n, host_.row_ptr.data(), host_.col.data(),
                           host_.weights.empty() ? nullptr : host_.weights.data(), (int)num_compute, (int)num_memory,
                           mirror_.empty() ? nullptr : mirror_.data(), &graph_),
        "partitioning the graph");
}

DistributedGraph::~DistributedGraph() { grd_graph_free(graph_); }

std::vector<GNode> DistributedGraph::translationTable(uint32_t c, uint32_t m) const
{
  uint64_t cnt = 0;
  check(grd_graph_table(graph_, (int)c, (int)m, nullptr, &cnt), "translation table");
  std::vector<GNode> out(cnt);
  check(grd_graph_table(graph_, (int)c, (int)m, out.data(), &cnt), "translation table");
  return out;
}

uint64_t DistributedGraph::numProxies() const
{
  uint64_t info[16];
  check(grd_graph_info(graph_, info), "graph info");
  return info[5];
}

uint64_t DistributedGraph::numCrossMirrors() const
{
  uint64_t info[16];
  check(grd_graph_info(graph_, info), "graph info");
  return info[6];
}
