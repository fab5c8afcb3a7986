// Exchange.hpp -- the on-device replacement of MPICore (reference
// include/MPI.hpp:8-124).
//
// Every compute and memory partition lives in one GPU's HBM, so the
// point-to-point messages of the reference become device kernels over wire
// buffers (grd_broadcast / grd_reduce) and the scalar allreduces become
// device reductions read once per superstep.  The verbs the operator API uses
// are kept: rank/size, barrier, allReduce on host scalars (single process:
// identity), and the bytes-moved counter.
#ifndef GRUDON_EXCHANGE_HPP
#define GRUDON_EXCHANGE_HPP

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "grudon_hip.h"

class Exchange
{
 public:
  explicit Exchange(int device = 0)
  {
    if (grd_create(device, &ctx_) != GRD_OK)
      throw std::runtime_error(std::string("Exchange: cannot open the GPU: ") + grd_last_error());
  }
  Exchange(int /*argc*/, char ** /*argv*/) : Exchange(0) {}
  ~Exchange() { grd_destroy(ctx_); }
  Exchange(const Exchange &) = delete;
  Exchange &operator=(const Exchange &) = delete;

  int getRank() const { return 0; }
  int getNumProcs() const { return 1; }
  grd_ctx *context() const { return ctx_; }

  // every partition shares one HIP stream: ordering is the barrier
  void barrier() {}

  // one process: the reduction of a single contribution is the value
  template <typename T>
  void allReduce(const T *sendbuf, T *recvbuf, int count)
  {
    std::memcpy(recvbuf, sendbuf, sizeof(T) * (size_t)count);
  }

  uint64_t &getBytesMoved() { return bytes_moved_; }
  void incrementBytesMoved(uint64_t b) { bytes_moved_ += b; }
  void decrementBytesMoved(uint64_t b) { bytes_moved_ -= b; }

 private:
  grd_ctx *ctx_ = nullptr;
  uint64_t bytes_moved_ = 0;
};

#endif  // GRUDON_EXCHANGE_HPP
