// main.cpp -- command line of the operator path (reference src/main.cpp:16-148).
//
// Same flags as the reference (-g -c -m -t -p -h) plus the -a flag its README
// documents (pr | cc | sssp; prpull = normalised pull PageRank).  One process
// owns all C compute and M memory partitions on one GPU, so there is no
// mpirun and the node id is 0.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Exchange.hpp"
#include <type_traits>

#include "GraphAlgorithm.hpp"
#include "SelectorFile.hpp"
#include "graph_algorithms/cc.hpp"
#include "graph_algorithms/pr.hpp"
#include "graph_algorithms/sssp.hpp"

namespace {

const char *kHelp =
    "A Graph Processing Framework (MI355X, one process for every partition)\n"
    "Usage:\n"
    "  grudon [OPTION...]\n\n"
    "  -g, --graph-path arg           Input graph file (.gr, .mtx) or generator spec\n"
    "                                 rmat:SCALE[:EF[:SEED[:sym[:WSEED]]]] | er:N:DEG[:SEED[:WSEED]]\n"
    "  -c, --num-compute arg          Number of compute nodes\n"
    "  -m, --num-memory arg           Number of memory nodes\n"
    "  -t, --threads arg              Number of host threads for graph construction (default: 1)\n"
    "  -p, --partitioning-scheme arg  Partitioning scheme file (one memory partition per line)\n"
    "  -a, --algorithm arg            pr | cc | sssp | prpull (default: pr)\n"
    "  -h, --help                     Print usage\n"
    "      --sched arg                plain | heuristic | learned | ndp | fixed:N (default: plain)\n"
    "                                 ndp = the NDPEngine offload vote (edge shipping when\n"
    "                                 covered and small); fixed:4 = edge shipping\n"
    "      --selector arg             BF16 selector weights file for --sched learned\n"
    "      --source arg               SSSP source gid (default: 0)\n"
    "      --weighted                 SSSP: use the edge weights\n"
    "      --delta arg                SSSP: delta-stepping bucket width (default: 0 = off)\n"
    "      --max-iter arg             iteration cap (default: 1000)\n"
    "      --print-state              print every vertex value\n"
    "      --verify                   run the algorithm's invariant check\n"
    "      --write-gr arg             also write the input graph as a Galois .gr file\n"
    "      --value-bytes arg          vertex property width: 4 (default) or 8 (CC<uint64_t>,\n"
    "                                 SSSP<uint64_t>; PageRank<double> is refused)\n"
    "      --result arg               write the result by gid (raw values of the property width)\n";

struct Options {
  std::string graph_path, pfile, algo = "pr", sched = "plain", selector, write_gr, result;
  size_t num_compute = 0, num_memory = 0, threads = 1;
  uint32_t source = 0, delta = 0, max_iter = 1000, value_bytes = 4;
  bool weighted = false, print_state = false, verify = false, help = false;
  bool has_g = false, has_c = false, has_m = false;
};

Options parse(int argc, char **argv)
{
  Options o;
  auto need = [&](int &i, const std::string &name) -> std::string {
    if (i + 1 >= argc) throw std::runtime_error("Option '" + name + "' is missing an argument");
    return argv[++i];
  };
  auto num = [](const std::string &s, const std::string &name) -> unsigned long long {
    char *end = nullptr;
    unsigned long long v = std::strtoull(s.c_str(), &end, 10);
    if (s.empty() || *end || s[0] == '-') throw std::runtime_error("Argument '" + s + "' failed to parse for option " + name);
    return v;
  };
  auto num32 = [&](const std::string &s, const std::string &name) -> uint32_t {
    const unsigned long long v = num(s, name);
    if (v > 0xFFFFFFFFull) throw std::runtime_error("Argument '" + s + "' failed to parse for option " + name);
    return (uint32_t)v;
  };
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "-g" || a == "--graph-path") o.graph_path = need(i, a), o.has_g = true;
    else if (a == "-c" || a == "--num-compute") o.num_compute = num(need(i, a), a), o.has_c = true;
    else if (a == "-m" || a == "--num-memory") o.num_memory = num(need(i, a), a), o.has_m = true;
    else if (a == "-t" || a == "--threads") o.threads = num(need(i, a), a);
    else if (a == "-p" || a == "--partitioning-scheme") o.pfile = need(i, a);
    else if (a == "-a" || a == "--algorithm") o.algo = need(i, a);
    else if (a == "-h" || a == "--help") o.help = true;
    else if (a == "--sched") o.sched = need(i, a);
    else if (a == "--selector") o.selector = need(i, a);
    else if (a == "--source") o.source = num32(need(i, a), a);
    else if (a == "--weighted") o.weighted = true;
    else if (a == "--delta") o.delta = num32(need(i, a), a);
    else if (a == "--max-iter") o.max_iter = num32(need(i, a), a);
    else if (a == "--print-state") o.print_state = true;
    else if (a == "--verify") o.verify = true;
    else if (a == "--write-gr") o.write_gr = need(i, a);
    else if (a == "--result") o.result = need(i, a);
    else if (a == "--value-bytes") o.value_bytes = num32(need(i, a), a);
    else throw std::runtime_error("Option '" + a + "' does not exist");
  }
  return o;
}

template <typename VP>
int run_algorithm(GraphAlgorithm<VP> &alg, const Options &o)
{
  alg.init();
  alg.run();
  if (o.print_state) alg.printState();
  if (o.verify) alg.verify();
  if (!o.result.empty()) {
    std::vector<VP> r = alg.result();
    std::ofstream out(o.result, std::ios::binary);
    out.write(reinterpret_cast<const char *>(r.data()), (std::streamsize)(r.size() * sizeof(VP)));
    if (!out) throw std::runtime_error("cannot write " + o.result);
  }
  return 0;
}

template <typename VP>
void write_input_gr(GraphAlgorithm<VP> &alg, const Options &o)
{
  if (o.write_gr.empty()) return;
  const HostGraph &g = alg.graph().host();
  if (grd_graph_write_gr(o.write_gr.c_str(), g.n, g.row_ptr.data(), g.col.data(),
                         g.weights.empty() ? nullptr : g.weights.data()) != GRD_OK)
    throw std::runtime_error(grd_last_error());
}

template <typename VP>
int run_cc(const Options &o, Exchange &net, grd_sched sched, int fixed, uint32_t node_id, std::string &pfile)
{
  CC<VP> alg(COMPUTE_NODE, "", o.graph_path, o.num_compute, o.num_memory, node_id, net, pfile);
  alg.setSchedule(sched, fixed);
  alg.setMaxIterations(o.max_iter);
  write_input_gr(alg, o);
  return run_algorithm(alg, o);
}

template <typename VP>
int run_sssp(const Options &o, Exchange &net, grd_sched sched, int fixed, uint32_t node_id, std::string &pfile)
{
  SSSP<VP> alg(COMPUTE_NODE, "", o.graph_path, o.num_compute, o.num_memory, node_id, net, pfile);
  alg.setSchedule(sched, fixed);
  alg.setMaxIterations(o.max_iter);
  alg.setSource(o.source);
  alg.setWeighted(o.weighted);
  alg.setDelta(o.delta);
  write_input_gr(alg, o);
  return run_algorithm(alg, o);
}

template <typename VP>
int run_pr(const Options &o, Exchange &net, grd_sched sched, int fixed, uint32_t node_id, std::string &pfile)
{
  std::unique_ptr<PageRank<VP>> alg;
  if constexpr (std::is_same<VP, float>::value) {
    if (o.algo == "prpull")
      alg.reset(new PageRankPull<VP>(COMPUTE_NODE, "", o.graph_path, o.num_compute, o.num_memory, node_id, net, pfile));
  }
  if (!alg)
    alg.reset(new PageRank<VP>(COMPUTE_NODE, "", o.graph_path, o.num_compute, o.num_memory, node_id, net, pfile));
  alg->setSchedule(sched, fixed);
  alg->setMaxIterations(o.max_iter);
  write_input_gr(*alg, o);
  return run_algorithm(*alg, o);
}

}  // namespace

int main(int argc, char **argv)
{
  Options o;
  try {
    o = parse(argc, argv);
    if (o.help) {
      std::cout << kHelp << std::endl;
      return 0;
    }
    if (!o.has_g) {
      std::cerr << "Please provide a graph path" << std::endl;
      return 1;
    }
    if (!o.has_c) {
      std::cerr << "Please provide the number of master nodes" << std::endl;
      return 1;
    }
    if (!o.has_m) {
      std::cerr << "Please provide the number of mirror nodes" << std::endl;
      return 1;
    }
  } catch (const std::exception &e) {
    std::cerr << "Error parsing options: " << e.what() << std::endl;
    return 1;
  }
  if (o.algo != "pr" && o.algo != "cc" && o.algo != "sssp" && o.algo != "prpull") {
    std::cerr << "Error parsing options: unknown algorithm '" << o.algo << "' (pr | cc | sssp | prpull)" << std::endl;
    return 1;
  }
  if (o.value_bytes != 4 && o.value_bytes != 8) {
    std::cerr << "Error parsing options: --value-bytes must be 4 or 8" << std::endl;
    return 1;
  }
  if (o.num_compute < 1 || o.num_memory < 1) {
    std::cerr << "Error parsing options: -c and -m must be at least 1" << std::endl;
    return 1;
  }
  grd_sched sched = GRD_SCHED_PLAIN;
  int fixed = 0;
  if (o.sched == "heuristic") sched = GRD_SCHED_HEURISTIC;
  else if (o.sched == "learned") sched = GRD_SCHED_LEARNED;
  else if (o.sched == "ndp") sched = GRD_SCHED_NDP_ENGINE;
  else if (o.sched.rfind("fixed:", 0) == 0) {
    sched = GRD_SCHED_FIXED;
    const std::string n = o.sched.substr(6);
    if (n.size() != 1 || n[0] < '0' || n[0] > '4') {
      std::cerr << "Error parsing options: fixed schedule option must be 0..4, not '" << n << "'" << std::endl;
      return 1;
    }
    fixed = n[0] - '0';
  } else if (o.sched != "plain") {
    std::cerr << "Error parsing options: unknown schedule '" << o.sched << "'" << std::endl;
    return 1;
  }

  try {
    Exchange net(0);
    grd_set_host_threads(net.context(), (int)o.threads);
    std::printf("[%s] Graph Path: %s\n[%s] Number of Compute Nodes: %zu\n[%s] Number of Memory Nodes: %zu\n"
                "[%s] Number of Threads: %zu\n", "info", o.graph_path.c_str(), "info", o.num_compute, "info",
                o.num_memory, "info", o.threads);
    if (!o.selector.empty()) {
      const std::vector<uint16_t> w = read_selector_weights(o.selector);
      if (grd_selector_load(net.context(), w.data(), w.size(), GRD_SELECTOR_MARGIN) != GRD_OK)
        throw std::runtime_error(std::string("selector: ") + grd_last_error());
    }
    const uint32_t node_id = 0;
    std::string pfile = o.pfile;
    const bool wide = o.value_bytes == 8;
    if (o.algo == "cc")
      return wide ? run_cc<uint64_t>(o, net, sched, fixed, node_id, pfile)
                  : run_cc<uint32_t>(o, net, sched, fixed, node_id, pfile);
    if (o.algo == "sssp")
      return wide ? run_sssp<uint64_t>(o, net, sched, fixed, node_id, pfile)
                  : run_sssp<uint32_t>(o, net, sched, fixed, node_id, pfile);
    if (o.algo == "prpull" && wide) throw std::runtime_error("the pull PageRank keeps 4-byte (float) ranks");
    return wide ? run_pr<double>(o, net, sched, fixed, node_id, pfile)
                : run_pr<float>(o, net, sched, fixed, node_id, pfile);
  } catch (const std::exception &e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 2;
  }
}
