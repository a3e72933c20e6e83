"""Independent pure-Python restatements used to pin the C oracle.

These are written from the mathematics (polynomial arithmetic over GF(2),
the BLAKE3 specification, SplitMix64), not from the oracle's code, so a
mistake in one is caught by disagreement with the other.  Small inputs only.
"""

GHASH_MOD = (1 << 128) | 0x87


def f128_from(lo, hi):
    return lo | (hi << 64)


def gf_mul(a, b, mod=GHASH_MOD, deg=128):
    """Carry-less product of Python ints reduced mod `mod` (degree `deg`)."""
    p = 0
    while b:
        if b & 1:
            p ^= a
        a <<= 1
        b >>= 1
    for i in range(p.bit_length() - 1, deg - 1, -1):
        if (p >> i) & 1:
            p ^= mod << (i - deg)
    return p


def gf_pow(a, e, mod=GHASH_MOD, deg=128):
    r = 1
    while e:
        if e & 1:
            r = gf_mul(r, a, mod, deg)
        a = gf_mul(a, a, mod, deg)
        e >>= 1
    return r


def gf_inv(a):
    return gf_pow(a, (1 << 128) - 2)


def gf8_mul(a, b):
    return gf_mul(a, b, 0x11B, 8)


# --------------------------------------------------------------------------- BLAKE3
IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
M32 = 0xFFFFFFFF
CHUNK_START, CHUNK_END, PARENT, ROOT = 1, 2, 4, 8


def _rotr(x, n):
    return ((x >> n) | (x << (32 - n))) & M32


def compress(cv, block, counter, block_len, flags):
    v = list(cv) + IV[:4] + [counter & M32, (counter >> 32) & M32, block_len, flags]
    m = list(block)

    def g(a, b, c, d, x, y):
        v[a] = (v[a] + v[b] + x) & M32
        v[d] = _rotr(v[d] ^ v[a], 16)
        v[c] = (v[c] + v[d]) & M32
        v[b] = _rotr(v[b] ^ v[c], 12)
        v[a] = (v[a] + v[b] + y) & M32
        v[d] = _rotr(v[d] ^ v[a], 8)
        v[c] = (v[c] + v[d]) & M32
        v[b] = _rotr(v[b] ^ v[c], 7)

    for r in range(7):
        g(0, 4, 8, 12, m[0], m[1]); g(1, 5, 9, 13, m[2], m[3])
        g(2, 6, 10, 14, m[4], m[5]); g(3, 7, 11, 15, m[6], m[7])
        g(0, 5, 10, 15, m[8], m[9]); g(1, 6, 11, 12, m[10], m[11])
        g(2, 7, 8, 13, m[12], m[13]); g(3, 4, 9, 14, m[14], m[15])
        m = [m[PERM[i]] for i in range(16)]
    return [v[i] ^ v[i + 8] for i in range(8)] + [v[i + 8] ^ cv[i] for i in range(8)]


def _words(b):
    b = b + bytes(64 - len(b))
    return [int.from_bytes(b[4 * i:4 * i + 4], "little") for i in range(16)]


def _wbytes(w):
    return b"".join(x.to_bytes(4, "little") for x in w)


def _chunk_nodes(data):
    """Return the list of (cv, block_words, counter, block_len, flags) final nodes per chunk."""
    chunks = [data[i:i + 1024] for i in range(0, len(data), 1024)] or [b""]
    out = []
    for ci, ch in enumerate(chunks):
        cv = IV
        blocks = [ch[i:i + 64] for i in range(0, len(ch), 64)] or [b""]
        for bi, bl in enumerate(blocks):
            flags = (CHUNK_START if bi == 0 else 0) | (CHUNK_END if bi == len(blocks) - 1 else 0)
            if bi == len(blocks) - 1:
                out.append((cv, _words(bl), ci, len(bl), flags))
            else:
                cv = compress(cv, _words(bl), ci, 64, flags)[:8]
    return out


def _root_node(data):
    nodes = _chunk_nodes(data)
    # Build the left-heavy binary tree: left subtree = largest power of two < n chunks.
    def build(lo, hi):
        n = hi - lo
        if n == 1:
            return nodes[lo]
        half = 1
        while half * 2 < n:
            half *= 2
        lcv = _node_cv(build(lo, lo + half))
        rcv = _node_cv(build(lo + half, hi))
        return (IV, lcv + rcv, 0, 64, PARENT)
    return build(0, len(nodes))


def _node_cv(node):
    cv, bw, ctr, bl, fl = node
    return compress(cv, bw, ctr, bl, fl)[:8]


def blake3(data, out_len=32):
    cv, bw, _ctr, bl, fl = _root_node(data)
    out = b""
    k = 0
    while len(out) < out_len:
        out += _wbytes(compress(cv, bw, k, bl, fl | ROOT))
        k += 1
    return out[:out_len]


def blake3_non_root_cv(data):
    return _wbytes(_node_cv(_root_node(data)))


def blake3_parent_cv(left, right):
    return _wbytes(compress(IV, _words(left + right), 0, 64, PARENT)[:8])


# --------------------------------------------------------------------------- SplitMix64
def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)
