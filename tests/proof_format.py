"""Field map of an R1csProofBundleLigerito byte string ("FLOCK" 04 02 +
bincode 1.x fixint, crates/flock-core/src/proof_io.rs:161-208; struct order of
proof.rs:18-22, zerocheck.rs:217-231, lincheck.rs:436-447, pcs.rs:51,
ring_switch.rs:2503, pcs/ligerito.rs:1504-1545).

`fields(b)` lists (name, offset, length) for every leaf field in order;
`first_difference(a, b)` names the field holding the first differing byte,
so a parity failure says which protocol step diverged.
"""
import struct


class _Reader:
    def __init__(self, b):
        self.b = b
        self.o = 0
        self.out = []

    def take(self, name, n):
        if self.o + n > len(self.b):
            raise ValueError(f"truncated at {name} (offset {self.o}, need {n})")
        self.out.append((name, self.o, n))
        v = self.b[self.o:self.o + n]
        self.o += n
        return v

    def u64(self, name):
        return struct.unpack("<Q", self.take(name, 8))[0]

    def u32(self, name):
        return struct.unpack("<I", self.take(name, 4))[0]

    def fvec(self, name):
        n = self.u64(name + ".len")
        self.take(name, 16 * n)
        return n

    def opening(self, name):
        rows = self.u64(name + ".rows.len")
        for i in range(rows):
            self.fvec(f"{name}.rows[{i}]")
        h = self.u64(name + ".hashes.len")
        self.take(name + ".hashes", 32 * h)


def fields(b):
    r = _Reader(bytes(b))
    r.take("magic", 5)
    r.take("version", 2)
    r.take("commitment.root", 32)
    for k in ("m", "log_inv_rate", "log_batch_size"):
        r.u64("commitment." + k)
    r.u32("commitment.profile")
    r.u32("commitment.merkle_hash")
    r.fvec("zerocheck.round1_ab")
    r.fvec("zerocheck.round1_c")
    n = r.u64("zerocheck.rounds.len")
    for i in range(n):
        r.take(f"zerocheck.rounds[{i}]", 32)
    for k in ("a", "b", "c"):
        r.take("zerocheck.final_" + k, 16)
    n = r.u64("lincheck.rounds.len")
    for i in range(n):
        r.take(f"lincheck.rounds[{i}]", 32)
    r.fvec("lincheck.z_partial")
    n = r.u64("ring_switch.s_hat_v.len")
    for i in range(n):
        r.fvec(f"ring_switch.s_hat_v[{i}]")
    r.take("ligerito.initial_root", 32)
    r.opening("ligerito.initial_proof")
    n = r.u64("ligerito.roots.len")
    for i in range(n):
        r.take(f"ligerito.roots[{i}]", 32)
    n = r.u64("ligerito.recursive_proofs.len")
    for i in range(n):
        r.opening(f"ligerito.recursive_proofs[{i}]")
    r.fvec("ligerito.final.yr")
    r.opening("ligerito.final.opening")
    n = r.u64("ligerito.sumcheck_msgs.len")
    for i in range(n):
        r.take(f"ligerito.sumcheck_msgs[{i}]", 32)
    n = r.u64("ligerito.nonces.len")
    r.take("ligerito.nonces", 8 * n)
    r.fvec("ligerito.ood")
    n = r.u64("ligerito.fold_nonces.len")
    r.take("ligerito.fold_nonces", 8 * n)
    if r.o != len(r.b):
        raise ValueError(f"{len(r.b) - r.o} trailing bytes")
    return r.out


def first_difference(a, b):
    """None when equal, else a short description of the first differing field."""
    if a == b:
        return None
    n = min(len(a), len(b))
    off = next((i for i in range(n) if a[i] != b[i]), n)
    try:
        fa = fields(a)
    except ValueError as e:
        return f"first differing byte {off} (lengths {len(a)} / {len(b)}); unparsable: {e}"
    name = next((nm for nm, o, ln in fa if o <= off < o + ln), "<end>")
    return f"first differing byte {off} in {name} (lengths {len(a)} / {len(b)})"
