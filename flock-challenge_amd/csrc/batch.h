// Batched proving with the fronts fused across instances (batch.cpp).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

// One proof of a batch: device compressions (blocks, n_blocks) or, when
// blocks is null, generate_compressions(log2, seed) expanded on the device.
struct BatchItem {
  const flock_compression *blocks = nullptr;
  uint64_t n_blocks = 0;
  uint32_t log2 = 0;
  uint64_t seed = 0;
};

struct BatchStats {
  double front_s = 0;           // the fused fronts, seconds
  bool graph_replayed = false;  // the fronts replayed a captured graph
};

// Proof bytes of every item (the same as a solo prove of its input), the
// fronts fused across items (replayed from a HIP graph when use_graph and
// a batch of the same shape ran before), the rest on n_workers worker
// contexts.  done_s[i] = seconds from the call's start to proof i's bytes.
int prove_batch(Ctx *ctx, const std::vector<BatchItem> &items, const uint8_t *domain, size_t domain_len,
                int n_workers, bool use_graph, std::vector<std::vector<uint8_t>> &proofs, std::vector<double> &done_s,
                BatchStats *stats);

}  // namespace flock
