// Graph.hpp -- node roles and offload decisions of the operator API
// (reference include/Graph.hpp:9-24).  One process owns every partition, so
// GNode is a 32-bit vertex id and the graph type itself lives behind the
// C-ABI (grd_graph).
#ifndef GRUDON_GRAPH_HPP
#define GRUDON_GRAPH_HPP

#include <cstdint>

using GNode = uint32_t;

enum NODE_TYPE
{
  COMPUTE_NODE,
  MEMORY_NODE,
  SWITCH_NODE
};

enum OFFLOAD_DECISION
{
  NDP_OFFLOAD,
  NO_OFFLOAD,
  INC_OFFLOAD,
};

#endif  // GRUDON_GRAPH_HPP
