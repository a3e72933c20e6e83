// Batched proving with the fronts fused across instances (batch.cpp).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "prover.h"

namespace flock {

// One proof of a batch: device compressions (blocks, n_blocks) or, when
// blocks is null, generate_compressions(log2, seed) expanded on the device.
struct BatchItem {
  const flock_compression *blocks = nullptr;
  uint64_t n_blocks = 0;
  uint32_t log2 = 0;
  uint64_t seed = 0;
};

// One instance of a fused front: its compressions (device) and their count,
// 2^n_log >= n_real, the buffers the front fills (prover.h ProofFront) and,
// once it has run, the commitment root.  m and k are filled in.
struct FrontJob {
  const flock_compression *blocks = nullptr;
  uint64_t n_real = 0;
  int n_log = 0, m = 0, k = 0;
  ProofFront fb;
  uint8_t root[32];
};
// Captured front graphs per batch shape.
struct FrontGraphs;
// The fronts of every job, fused (one launch per step), on ctx's stream;
// with graphs, the launch sequence is captured per shape and replayed.
// Waits for the roots.
int run_fronts(Ctx *ctx, std::vector<FrontJob> &jobs, FrontGraphs *graphs, bool *replayed);

struct BatchStats {
  double front_s = 0;           // the fused fronts, seconds
  bool graph_replayed = false;  // the fronts replayed a captured graph
  size_t fused = 0;             // items whose fronts ran fused
};

// Items of at most 2^FUSE_MAX_LOG2 compressions have their fronts fused:
// their launches are short enough for launch overhead to matter.  Larger
// ones run as whole proofs on the worker streams, in parallel with the
// fused fronts.
constexpr int FUSE_MAX_LOG2 = 14;

// Proof bytes of every item (the same as a solo prove of its input).  The
// fronts of the items of at most 2^FUSE_MAX_LOG2 compressions run fused on
// ctx's stream (replayed from a HIP graph when use_graph and a batch with
// the same small items ran before), then the n_workers worker contexts take
// larger items as whole proofs (largest first) and the rest of the small
// items' proofs, one worker at a time on small items while large ones
// remain.  done_s[i] = seconds from the call's start to proof i's bytes.
int prove_batch(Ctx *ctx, const std::vector<BatchItem> &items, const uint8_t *domain, size_t domain_len,
                int n_workers, bool use_graph, std::vector<std::vector<uint8_t>> &proofs, std::vector<double> &done_s,
                BatchStats *stats);

}  // namespace flock
