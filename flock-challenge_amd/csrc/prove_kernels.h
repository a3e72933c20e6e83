// Device building blocks of the prover (one launcher per kernel family).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

int witness_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_real, int n_log, void *d_z,
                   void *d_a, void *d_b);
// One instance of a fused witness launch: its compressions, their count and
// the padded count 2^n_log, its packed z, a, b, and its first workgroup
// (64 compressions per workgroup) in the launch.
struct WitnessJob {
  const flock_compression *blocks;
  uint64_t n_real, n_total;
  uint64_t *z, *a, *b;
  uint64_t wg_begin;
};
// Every job's witness in one launch of total_wgs workgroups (d_jobs in
// device memory, sorted by wg_begin).
int witness_batch_device(Ctx *ctx, hipStream_t st, const WitnessJob *d_jobs, uint32_t n_jobs, uint64_t total_wgs);

}  // namespace flock

namespace flock {
size_t zc_tables_bytes();
// Write seq to the mapped host word d_flag after the stream's earlier work.
int signal_device(Ctx *ctx, hipStream_t st, void *d_flag, uint32_t seq);
// A zerocheck round's transcript step (device transcript), run by the
// kernel that reduces the round's message: absorb the message (G(1), G(inf))
// into the transcript state and sample the round's challenge into rho (both
// device memory, transcript_dev.h).  Null state: no step.
struct FsStep {
  void *state = nullptr;
  void *rho = nullptr;
  void *digest = nullptr;  // observe-only steps: the state's digest then goes here (32 bytes)
};
// out[w] = xor_r in[r * width + w] (width <= 4); a nonzero seq is then
// published at the context's readback flag (see Ctx::READBACK_FLAG), which
// the reductions ending in the readback area use in place of signal_device.
// With fs.state the kernel then runs the transcript step over out[0],
// out[1]: observe both and sample into fs.rho (width 2), or, with fs.rho
// null, observe them only (width 2 or 3).
int xor_rows_device(Ctx *ctx, hipStream_t st, const void *in, uint64_t count, uint32_t width, void *out,
                    uint32_t seq = 0, const FsStep &fs = FsStep());
// Zerocheck round 1, product part, in two steps: zc_y_device computes
// Y_j(x_mid, x_hi) from the rows (no challenge involved), zc_round1_device
// then d_q[4 x_mid + q4][8 j + t] (64 x 128 F128) = sum_{x_hi} Ehi(x_hi)
// bit_t(Y_{16 q4 + j}(x_mid, x_hi)); P^AB follows on the host.
int zc_y_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, int m, const void *d_tables);
int zc_round1_device(Ctx *ctx, hipStream_t st, int m, const void *d_ehi, void *d_q);
// Q_k[il][r] = sum_ih bit_r(elem[il + 2^il_log ih]) hi_k[ih] over 128-bit
// elements (2^il_log x n_ih of them); hi1 may be null (one claim).  out:
// [2^(il_log-4)][n_claims][16][128] F128.
int bitfold_device(Ctx *ctx, hipStream_t st, const void *elems, int il_log, uint64_t n_ih, const void *hi0,
                   const void *hi1, void *out);
// out[y][w] = xor_x in[y][x][w] (ny stacks of nx rows of width F128).
int xor_reduce_device(Ctx *ctx, hipStream_t st, const void *in, uint32_t ny, uint32_t nx, uint32_t width, void *out);
int eq_expand_device(Ctx *ctx, hipStream_t st, const void *lo, int k, const void *hi, int n_hi_log, void *out);
// (G(1), G(inf)) of one multilinear zerocheck round with the eq table split
// as Elo (x) Ehi, Elo over the low ZS_LOG variables (2^ZS_LOG entries), Ehi
// over the rest (half >> ZS_LOG entries, at least 2); writes Ehi folded for
// the next round into Ehi_next (half >> (ZS_LOG + 1) entries) unless it is
// null.  Here and in the other reductions below, a nonzero seq is published
// once the result (out2 / out3) is written (xor_rows_device).
constexpr int ZS_LOG = 10;
int zc_round_split_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, const void *Elo, const void *Ehi,
                          uint64_t half, void *out2, void *Ehi_next, uint32_t seq = 0, const FsStep &fs = FsStep());
// Round 0 split at ZS_LOG + 1 over the packed rows a, b (A = L(row) through
// the 256-entry nibble tables d_lw, entry v * 16 + p): Elo has 2^(ZS_LOG + 1)
// entries, Ehi half >> (ZS_LOG + 1) (at least 2); Ehi is not folded.
int zc_round0_split_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, const void *d_lw, const void *Elo,
                           const void *Ehi, uint64_t half, void *out2, uint32_t seq = 0,
                           const FsStep &fs = FsStep());
// The fold after zerocheck round 0, reading the packed rows a, b and
// evaluating A = L(row) through 2048-entry byte tables (entry v * 8 + p):
// d_lw8 of L and d_lwr8 of rho L.
int zc_fold0_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, const void *d_lw8, const void *d_lwr8,
                    uint64_t half, void *A2, void *B2);
// d_r (may be null): read the challenge from device memory instead of r.
int fold_pair_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, uint64_t half, F128 r, void *A2, void *B2,
                     const void *d_r = nullptr);
// A round fused with the fold before it: fold the previous round's tables
// (4 half entries) by its challenge into A2, B2 (2 half), compute the new
// round's message over them and publish it (seq) as xor_rows_device does.
// Up to SMALL_PAIRS pairs this is one single-workgroup launch (reduction
// included), above it a grid plus the reduction.  Zerocheck: out2 =
// (G(1), G(inf)) with the whole eq table E (half entries) and the next
// round's eq table (E folded) into E_next unless it is null; Ligerito: out3 = (u0, u2, 0) as lig_msg_device without the
// evaluation.
constexpr uint64_t SMALL_PAIRS = 256;
int zc_fold_round_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, F128 rho, const void *E,
                         uint64_t half, void *A2, void *B2, void *E_next, void *out2, uint32_t seq,
                         const void *d_rho = nullptr, const FsStep &fs = FsStep());
// With d_r the Ligerito fold reads its challenge from device memory, and
// with fs.state (fs.rho null) the message's reduction then observes (u0, u2)
// in the device transcript.
int lig_fold_msg_device(Ctx *ctx, hipStream_t st, const void *f, const void *b, F128 r, uint64_t half, void *f2,
                        void *b2, void *out3, uint32_t seq, const void *d_r = nullptr, const FsStep &fs = FsStep());
int lig_msg_device(Ctx *ctx, hipStream_t st, const void *f, const void *b, uint64_t half, int with_eval, void *out3,
                   uint32_t seq = 0, const FsStep &fs = FsStep());
int axpy_device(Ctx *ctx, hipStream_t st, void *acc, const void *x, uint64_t n, F128 beta);
// A transposed R1CS matrix on the device for the lincheck combination: the
// nonzero rows of column c are rows[ptr[c] .. ptr[c + 1]) (16-bit, ascending),
// those below row 2^13 up to split[c]; rows is padded by 8 entries.
struct LcCsc {
  const uint32_t *ptr, *split;
  const uint16_t *rows;
};
int lc_comb_device(Ctx *ctx, hipStream_t st, LcCsc a, LcCsc b, const void *eq, F128 alpha, F128 beta, uint32_t pin,
                   void *comb);
// out[i] = M_0(lo0[i mod 2^k] hi0[i >> k]) + M_1(lo1[i mod 2^k] hi1[i >> k])
// for i < 2^(k + n_hi_log), M_c given by its nibble tables (tables[c][p][v]
// = M_c(v x^(4p)), 2 x 32 x 16 F128).
int rs_basis_device(Ctx *ctx, hipStream_t st, const void *tables, const void *lo0, const void *lo1, int k,
                    const void *hi0, const void *hi1, int n_hi_log, void *out);
// Whole forward (layers l0..k-1) or transposed (k-1..l0) NTT over 2^k
// positions x 2^log_nl lanes, up to 4 layers per launch.
int ntt_narrow_device(Ctx *ctx, hipStream_t st, void *d, const void *tw, int k, int l0, int log_nl, int transposed);
// out (device-visible) = *best, then *best = all-ones for the next search
// (pow_chain_device); then publishes a nonzero seq as xor_rows_device does.
int pow_take_device(Ctx *ctx, hipStream_t st, void *d_best, void *out, uint32_t seq = 0);
int gather_device(Ctx *ctx, hipStream_t st, const void *src, const uint64_t *d_idx, uint32_t n, uint32_t unit_bytes,
                  void *out);
int scatter_device(Ctx *ctx, hipStream_t st, void *dst, const uint64_t *d_idx, const void *d_val, uint32_t n);
int fold_replicas_device(Ctx *ctx, hipStream_t st, const void *in, uint64_t n, uint32_t reps, void *out);
}  // namespace flock
