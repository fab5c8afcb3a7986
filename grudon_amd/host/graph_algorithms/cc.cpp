// cc.cpp -- CC operators bound to the device path (reference
// src/graph_algorithms/cc.cpp:18-177).
#include "graph_algorithms/cc.hpp"

#include <cstdio>

template <typename VP>
CC<VP>::CC(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
           size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file)
    : GraphAlgorithm<VP>(node_type, algorithm_name, graph_path, num_compute, num_memory, node_id, net,
                         partitioning_scheme_file)
{
  this->algorithm_name = "CC";
  this->bindDevice(GRD_CC);
}

// cc.cpp:22-39: label = gid, update = MAX, every vertex in the frontier
template <typename VP>
void CC<VP>::init()
{
  this->check(grd_init(this->state, &this->params), "init");
}

// cc.cpp:42-51: nothing to apply
template <typename VP>
void CC<VP>::apply_updates()
{
}

// cc.cpp:53-84: atomicMin of the label along the local out-edges
template <typename VP>
void CC<VP>::gen_updates()
{
  this->check(grd_traverse(this->state, this->option), "gen_updates");
}

// cc.cpp:86-107: label = update where it improved, frontier = improved
template <typename VP>
void CC<VP>::update_frontier()
{
  this->check(grd_update(this->state, &this->frontier_count), "update_frontier");
}

// cc.cpp:109-113: minUpdate on the vertex with local id lid of the compute partition the
// worker serves
template <typename VP>
void CC<VP>::aggregate(GNode &lid, const VP &buffer_val)
{
  this->check(grd_aggregate_one(this->state, (int)this->aggregate_partition, lid, &buffer_val), "aggregate");
}

template <typename VP>
bool CC<VP>::termination_check()
{
  return this->frontier_count == 0;
}

template <typename VP>
void CC<VP>::printState()
{
  std::vector<VP> r = this->result();
  for (size_t v = 0; v < r.size(); v++) std::printf("[Proc 0] Vertex/Dist: %zu/%llu\n", v, (unsigned long long)r[v]);
}

// cc.cpp:139-177: every edge must join equal labels (symmetric input)
template <typename VP>
void CC<VP>::verify()
{
  std::vector<VP> r = this->result();
  const HostGraph &g = this->graph().host();
  violations = 0;
  for (uint32_t u = 0; u < g.n; u++)
    for (uint32_t x = g.row_ptr[u]; x < g.row_ptr[u + 1]; x++)
      if (r[u] != r[g.col[x]]) {
        if (violations < 10)
          std::printf("[Proc 0] Vertex %u and %u have different components: %llu and %llu\n", u, g.col[x],
                      (unsigned long long)r[u], (unsigned long long)r[g.col[x]]);
        violations++;
      }
  std::printf("CC verify: %llu edges join different labels\n", (unsigned long long)violations);
}

template class CC<uint32_t>;
template class CC<uint64_t>;
