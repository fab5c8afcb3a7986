// GraphAlgorithm.cpp -- construction and the BSP superstep loop
// (reference src/GraphAlgorithm.cpp:68-701).
#include "GraphAlgorithm.hpp"

#include <cstdio>
#include <stdexcept>
#include <string>

template <typename VP>
GraphAlgorithm<VP>::GraphAlgorithm(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path,
                                   size_t num_compute, size_t num_memory, uint32_t node_id, Exchange &net,
                                   std::string &partitioning_scheme_file)
    : algorithm_name(algorithm_name),
      graph_path(graph_path),
      node_type(node_type),
      net(net),
      num_compute((uint32_t)num_compute),
      num_memory((uint32_t)num_memory),
      num_workers((uint32_t)(num_compute + num_memory))
{
  if (num_compute < 1 || num_memory < 1) throw std::runtime_error("need at least one compute and one memory partition");
  HostGraph h = load_host_graph(net, graph_path);
  dgraph.reset(new DistributedGraph(net, h, num_compute, num_memory, partitioning_scheme_file));
  // one process holds both roles: node_id names a compute partition
  // (node_id < C) or a memory partition (node_id - C); the other role's
  // worker takes partition 0 (src/main.cpp:92-93)
  const uint32_t C = (uint32_t)num_compute;
  const uint32_t cnode = node_id < C ? node_id : 0;
  const uint32_t mnode = node_id >= C && node_id < C + num_memory ? node_id : C;
  worker.reset(new UpdateWorker<VP>(dgraph.get(), num_compute, num_memory, cnode, COMPUTE_NODE, net));
  traverser.reset(new TraverseWorker<VP>(dgraph.get(), num_compute, num_memory, mnode, MEMORY_NODE, net));
  aggregate_partition = cnode;
  params.max_iter = MAX_ITERATIONS;
  params.damping = 0.85;
  params.tolerance = 1e-6;
  params.sched = GRD_SCHED_PLAIN;
}

template <typename VP>
GraphAlgorithm<VP>::~GraphAlgorithm()
{
  grd_state_free(state);
}

template <typename VP>
void GraphAlgorithm<VP>::check(grd_status st, const char *what)
{
  if (st != GRD_OK) throw std::runtime_error(std::string(algorithm_name) + ": " + what + ": " + grd_last_error());
}

template <typename VP>
void GraphAlgorithm<VP>::bindDevice(int grd_algo)
{
  if (state) grd_state_free(state);
  state = nullptr;
  check(grd_state_create_typed(dgraph->handle(), grd_algo, (uint32_t)sizeof(VP), &state),
        "allocating the algorithm state");
}

template <typename VP>
void GraphAlgorithm<VP>::setSchedule(grd_sched sched, int fixed_option)
{
  params.sched = (uint32_t)sched;
  params.fixed_option = (uint32_t)fixed_option;
}

template <typename VP>
void GraphAlgorithm<VP>::run()
{
  uint32_t iteration = 0;
  uint32_t completion = 0;
  net.barrier();
  worker_completion_count = 0;
  while (worker_completion_count != num_compute && iteration < max_iterations) {
    worker_completion_count = 0;
    completion = 0;
    // decision point of the superstep (GraphAlgorithm.cpp:187-195)
    check(grd_select_option(state, &option), "schedule selection");
    // compute -> memory: frontier bitvectors + packed properties (:199-343)
    check(grd_broadcast(state, option), "broadcast");
    // memory partitions: gen_updates (:348-356)
    traverser->traverse(*this);
    // memory -> compute: dirty bitvectors + packed updates, aggregate() in
    // bulk on the device (:377-581)
    check(grd_reduce(state, option), "reduce");
    // compute partitions: update_frontier + apply_updates (:643-653)
    worker->update(*this);
    if (termination_check()) completion = 1;
    // every compute partition shares the device-wide frontier count
    uint32_t all = completion ? num_compute : 0;
    net.allReduce(&all, &worker_completion_count, 1);
    std::printf("[Proc %u] Iteration: %u, WCC: %u\n", worker->node_id, iteration++, worker_completion_count);
  }
  check(grd_state_stats(state, &last_stats), "statistics");
  net.incrementBytesMoved(last_stats.bytes_exchanged);
  std::printf("STAT, %s, Iterations, %u\n", algorithm_name.c_str(), iteration);
  std::printf("STAT, %s, TotalBytesMoved, %llu\n", algorithm_name.c_str(),
              (unsigned long long)last_stats.bytes_exchanged);
  std::fflush(stdout);
}

template <typename VP>
std::vector<VP> GraphAlgorithm<VP>::result()
{
  std::vector<VP> out(dgraph->getNumVertices());
  check(grd_download(state, out.data()), "download");
  return out;
}

template <typename VP>
void GraphAlgorithm<VP>::exchangeView(int dir, uint32_t c, uint32_t m, std::vector<GNode> &gids,
                                      std::vector<VP> &values)
{
  static_assert(sizeof(VP) == 4 || sizeof(VP) == 8, "the exchange packs 4- or 8-byte properties");
  uint64_t k = 0;
  check(grd_pair_exchange(state, dir, (int)c, (int)m, nullptr, nullptr, &k), "exchange read-back");
  gids.resize(k);
  values.resize(k);
  check(grd_pair_exchange(state, dir, (int)c, (int)m, gids.data(), values.data(), &k), "exchange read-back");
}

template class GraphAlgorithm<float>;
template class GraphAlgorithm<double>;
template class GraphAlgorithm<uint64_t>;
template class GraphAlgorithm<uint32_t>;
