/*
 * Oracle: PCS commit = replicate message, interleaved NTT from layer
 * log_inv_rate, BLAKE3 Merkle tree over positions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* pcs/commit.rs:250-320 (commit), 451-467 (replicate_message_fill),
 * 722-784 (cpu_transform_and_tree); PcsParams helpers at :54-90. */
int o_pcs_commit(const opcs_params *p, const of128 *z_packed, of128 *codeword, uint8_t *tree, uint8_t root[32]) {
    int log_msg = p->m - 7;
    int log_dim = log_msg - p->log_batch_size;
    int k_code = log_dim + p->log_inv_rate;
    if (log_dim < 0 || p->log_inv_rate < 1) return -1;
    size_t num_ntts = (size_t)1 << p->log_batch_size;
    size_t npos = (size_t)1 << k_code;
    size_t msg_len = (size_t)1 << log_msg;
    size_t cw_len = npos * num_ntts;
    of128 *cw = codeword ? codeword : (of128 *)malloc(sizeof(of128) * cw_len);
    uint8_t *tr = tree ? tree : (uint8_t *)malloc(32 * (2 * npos - 1));
    if (!cw || !tr) return -2;
    for (size_t off = 0; off < cw_len; off += msg_len) memcpy(cw + off, z_packed, sizeof(of128) * msg_len);
    ontt t;
    ontt_init(&t, k_code);
    ontt_forward_interleaved(&t, cw, cw_len, num_ntts, p->log_inv_rate);
    ontt_free(&t);
    omerkle_tree((const uint8_t *)cw, cw_len * sizeof(of128), npos, tr);
    memcpy(root, tr + 32 * (2 * npos - 2), 32);
    if (!codeword) free(cw);
    if (!tree) free(tr);
    return 0;
}
