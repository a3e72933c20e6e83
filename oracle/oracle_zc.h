/* Oracle-internal protocol declarations.  TEST INFRASTRUCTURE ONLY. */
#ifndef FLOCK_ORACLE_ZC_H
#define FLOCK_ORACLE_ZC_H
#include "oracle_proto.h"
#include "ligerito_configs.h"

typedef struct {
    of128 z;
    of128 *r;    /* m entries: the zerocheck eq point (r[6..] = r_rest) */
    of128 *rhos; /* m - 6 multilinear challenges */
} ozc_claim;

typedef struct {
    of128 r_inner_skip;
    of128 r_inner_rest[8];
    of128 w;
} olc_claim;

void o_zc_r_vector(ofs *t, int m, of128 *r);
int o_zerocheck_prove(const uint64_t *a, const uint64_t *b, const uint64_t *c, int m, ofs *t, ozc_proof *pf,
                      ozc_claim *cl);
int o_zerocheck_verify(const ozc_proof *pf, int m, ofs *t, ozc_claim *cl);

/* lincheck: x_inner_rest (8), x_outer (m - 14) */
int o_lincheck_prove(const uint64_t *z, int m, of128 z_skip, const of128 *x_inner_rest, const of128 *x_outer,
                     ofs *t, olc_proof *pf, olc_claim *cl);
int o_lincheck_verify(const olc_proof *pf, int m, of128 z_skip, const of128 *x_inner_rest, const of128 *x_outer,
                      of128 v_a, of128 v_b, ofs *t, olc_claim *cl);
void o_comb_vector(of128 alpha, const of128 *eq_inner, of128 *comb);

/* PCS open (ring switch + Ligerito) */
typedef struct {
    of128 z_skip;
    of128 x_full[64]; /* x_inner_rest ++ x_outer, length m - 6 */
    of128 value;
} oclaim;

int o_pcs_open(const uint64_t *z, const opcs_params *pp, const of128 *codeword, const uint8_t *tree,
               const oclaim *claims, int n_claims, ofs *t, oproof *pf);
int o_pcs_verify(const oproof *pf, const opcs_params *pp, const oclaim *claims, int n_claims, ofs *t);

const olig_config *o_lig_config(int m, int profile);
void o_build_eq_lsb(const of128 *r, int n, of128 *out);

#endif
