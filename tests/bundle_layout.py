"""The byte layout of a proof bundle, checked field by field against counts
fixed outside the prover: the struct declarations (zerocheck.rs:217-231,
lincheck.rs:436-447, ring_switch.rs:2503-2506, pcs/ligerito.rs:1504-1545,
3672-3675; bincode 1.x fixint, proof_io.rs:161-208) and the Ligerito ladders
(crates/flock-core/configs/ligerito/m26_fast.toml, m32_fast.toml).

Every Vec length prefix must sit where the declaration order puts it and
hold the count those sources fix, so serializing two fields of different
lengths in the wrong order fails, and so does a wrong count.  The counts
are written out below as data, from the TOML files, not computed by the
prover's code.
"""
import struct

# (queries, lanes of that level's codeword rows, ood samples, fold grinding
# bits, k folds) per Ligerito level, from the ladder files
LADDERS = {
    # m26_fast.toml: initial_k 6; levels (queries, log_num_interleaved,
    # ood_samples, fold_grinding_bits, k_recursive); final_block yr_log_n 4
    26: {"levels": [(218, 6, 0, 13, 6), (106, 3, 1, 8, 3), (71, 3, 1, 5, 3), (53, 3, 1, 2, 3)], "yr_log_n": 4},
    # m32_fast.toml
    32: {"levels": [(218, 6, 0, 19, 6), (106, 3, 1, 14, 3), (71, 3, 1, 11, 3), (53, 3, 1, 8, 3),
                    (43, 3, 1, 6, 3), (36, 3, 1, 4, 3)], "yr_log_n": 4},
}


def expected_counts(m):
    """Counts the ladder fixes.  Fold nonces: one per fold whose tapered
    grind (fold_grinding_bits - j, pcs/ligerito.rs:6370-6385) is positive.
    Sumcheck messages: the initial one, one per fold, one per OOD sample and
    one per induced basis (every level but the last)."""
    lv = LADDERS[m]["levels"]
    folds = sum(k for (_, _, _, _, k) in lv)
    ood = sum(o for (_, _, o, _, _) in lv)
    fold_nonces = sum(1 for (_, _, _, fb, k) in lv for j in range(k) if fb - j > 0)
    return {
        "zc_rounds": m - 6,
        "roots": len(lv) - 1,
        "recursive_proofs": len(lv) - 2,
        "sumcheck_msgs": 1 + folds + ood + (len(lv) - 1),
        "nonces": len(lv),
        "ood": ood,
        "fold_nonces": fold_nonces,
        "yr": 1 << LADDERS[m]["yr_log_n"],
    }


class Layout:
    def __init__(self, b):
        self.b, self.o = bytes(b), 0

    def raw(self, n):
        v = self.b[self.o:self.o + n]
        assert len(v) == n, f"truncated at offset {self.o}"
        self.o += n
        return v

    def u64(self):
        return struct.unpack("<Q", self.raw(8))[0]

    def u32(self):
        return struct.unpack("<I", self.raw(4))[0]

    def count(self, what, want):
        at = self.o
        got = self.u64()
        assert got == want, f"{what}: length prefix at offset {at} is {got}, expected {want}"
        return got

    def fvec(self, what, want):
        self.count(what, want)
        self.raw(16 * want)

    def opening(self, what, rows, lanes):
        self.count(what + ".opened_rows", rows)
        for i in range(rows):
            self.fvec(f"{what}.opened_rows[{i}]", lanes)
        h = self.u64()  # the multiproof's length depends on the positions
        assert 0 < h <= rows * 32, f"{what}.merkle_proof: {h} hashes for {rows} rows"
        self.raw(32 * h)


def check_layout(proof, m):
    """Walk the bundle in declaration order; raises AssertionError naming the
    first field whose prefix or value differs from the fixed layout."""
    e = expected_counts(m)
    lv = LADDERS[m]["levels"]
    L = Layout(proof)
    assert L.raw(5) == b"FLOCK" and L.raw(2) == b"\x04\x02"
    L.raw(32)  # commitment root
    assert (L.u64(), L.u64(), L.u64(), L.u32(), L.u32()) == (m, 1, 6, 0, 1), "PcsParams"
    L.fvec("zerocheck.round1_ab", 64)
    L.fvec("zerocheck.round1_c", 64)
    L.count("zerocheck.multilinear_rounds", e["zc_rounds"])
    L.raw(32 * e["zc_rounds"])
    L.raw(48)  # final a, b, c evaluations
    L.count("lincheck.rounds", 8)
    L.raw(32 * 8)
    L.fvec("lincheck.z_partial", 64)
    L.count("pcs_open.ring_switches", 2)
    L.fvec("ring_switches[0].s_hat_v", 128)
    L.fvec("ring_switches[1].s_hat_v", 128)
    L.raw(32)  # ligerito.initial_root
    L.opening("initial_proof", lv[0][0], 1 << lv[0][1])
    L.count("recursive_roots", e["roots"])
    L.raw(32 * e["roots"])
    L.count("recursive_proofs", e["recursive_proofs"])
    for i in range(e["recursive_proofs"]):
        L.opening(f"recursive_proofs[{i}]", lv[i + 1][0], 1 << lv[i + 1][1])
    L.fvec("final_proof.yr", e["yr"])
    L.opening("final_proof", lv[-1][0], 1 << lv[-1][1])
    L.count("sumcheck_transcript", e["sumcheck_msgs"])
    L.raw(32 * e["sumcheck_msgs"])
    L.count("grinding_nonces", e["nonces"])
    L.raw(8 * e["nonces"])
    L.fvec("ood_values", e["ood"])
    L.count("fold_grinding_nonces", e["fold_nonces"])
    L.raw(8 * e["fold_nonces"])
    assert L.o == len(L.b), f"{len(L.b) - L.o} trailing bytes"
