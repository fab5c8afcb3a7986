// sssp.cpp -- SSSP operators bound to the device path (reference
// src/graph_algorithms/sssp.cpp:18-189).
#include "graph_algorithms/sssp.hpp"

#include <cstdio>
#include <limits>

template <typename VP>
SSSP<VP>::SSSP(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
               size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file)
    : GraphAlgorithm<VP>(node_type, algorithm_name, graph_path, num_compute, num_memory, node_id, net,
                         partitioning_scheme_file)
{
  this->algorithm_name = "SSSP";
  this->bindDevice(GRD_SSSP);
}

// sssp.cpp:18-47: every distance MAX, the source 0 and in the frontier
template <typename VP>
void SSSP<VP>::init()
{
  this->check(grd_init(this->state, &this->params), "init");
}

template <typename VP>
void SSSP<VP>::apply_updates()
{
}

// sssp.cpp:60-90: atomicMin of dist + 1 (or dist + w) along the out-edges
template <typename VP>
void SSSP<VP>::gen_updates()
{
  this->check(grd_traverse(this->state, this->option), "gen_updates");
}

// sssp.cpp:92-113
template <typename VP>
void SSSP<VP>::update_frontier()
{
  this->check(grd_update(this->state, &this->frontier_count), "update_frontier");
}

// sssp.cpp:115-119: minUpdate on the vertex with local id lid of the compute partition the
// worker serves
template <typename VP>
void SSSP<VP>::aggregate(GNode &lid, const VP &buffer_val)
{
  this->check(grd_aggregate_one(this->state, (int)this->aggregate_partition, lid, &buffer_val), "aggregate");
}

template <typename VP>
bool SSSP<VP>::termination_check()
{
  return this->frontier_count == 0;
}

template <typename VP>
void SSSP<VP>::printState()
{
  std::vector<VP> r = this->result();
  for (size_t v = 0; v < r.size(); v++) std::printf("[Proc 0] Vertex/Dist: %zu/%llu\n", v, (unsigned long long)r[v]);
}

// sssp.cpp:144-189: no edge may offer a shorter path; max and visited count
template <typename VP>
void SSSP<VP>::verify()
{
  std::vector<VP> r = this->result();
  const HostGraph &g = this->graph().host();
  const bool weighted = this->params.weighted && !g.weights.empty();
  const VP inf = std::numeric_limits<VP>::max();
  violations = 0;
  uint64_t visited = 0;
  VP maxd = 0;
  for (uint32_t u = 0; u < g.n; u++) {
    VP sd = r[u];
    if (sd == inf) continue;
    for (uint32_t x = g.row_ptr[u]; x < g.row_ptr[u + 1]; x++) {
      uint64_t cand = (uint64_t)sd + (weighted ? g.weights[x] : 1u);
      if ((uint64_t)r[g.col[x]] > cand) {
        if (violations < 10)
          std::printf("Wrong label: %llu on node: %u, correct label from src node %u is %llu\n",
                      (unsigned long long)r[g.col[x]],
                      g.col[x], u, (unsigned long long)cand);
        violations++;
      }
    }
    visited++;
    if (sd > maxd) maxd = sd;
  }
  std::printf("Max Dist: %llu\nTotal Visited Dist: %llu\nSSSP verify: %llu violating edges\n",
              (unsigned long long)maxd,
              (unsigned long long)visited, (unsigned long long)violations);
}

template class SSSP<uint32_t>;
template class SSSP<uint64_t>;
