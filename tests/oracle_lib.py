"""ctypes bindings to the C oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class F128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


class F256(ctypes.Structure):
    _fields_ = [("r0", ctypes.c_uint64), ("r1", ctypes.c_uint64), ("r2", ctypes.c_uint64), ("r3", ctypes.c_uint64)]


class PcsParams(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("log_inv_rate", ctypes.c_int), ("log_batch_size", ctypes.c_int),
                ("profile", ctypes.c_int), ("merkle_hash", ctypes.c_int)]


class Ntt(ctypes.Structure):
    _fields_ = [("log_d", ctypes.c_int), ("twiddles", ctypes.c_void_p)]


P = ctypes.c_void_p
U8P = ctypes.POINTER(ctypes.c_uint8)
SZ = ctypes.c_size_t


def _declare(L):
    L.of128_mul.restype = F128
    L.of128_mul.argtypes = [F128, F128]
    L.of128_mul_portable.restype = F128
    L.of128_mul_portable.argtypes = [F128, F128]
    L.of128_inv.restype = F128
    L.of128_inv.argtypes = [F128]
    L.of128_mul_by_x.restype = F128
    L.of128_mul_by_x.argtypes = [F128]
    L.of128_mul_unreduced.restype = F256
    L.of128_mul_unreduced.argtypes = [F128, F128]
    L.of256_reduce.restype = F128
    L.of256_reduce.argtypes = [F256]
    L.of8_mul.restype = ctypes.c_uint8
    L.of8_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
    L.of8_inv.restype = ctypes.c_uint8
    L.of8_inv.argtypes = [ctypes.c_uint8]
    L.of_phi8.restype = F128
    L.of_phi8.argtypes = [ctypes.c_uint8]
    L.ob3_compress.argtypes = [P, P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, P]
    L.ob3_hash.argtypes = [P, SZ, P]
    L.ob3_leaf_cv.argtypes = [P, SZ, P]
    L.ob3_parent_cv.argtypes = [P, P, P]
    L.o_generate_compressions.argtypes = [ctypes.c_uint32, ctypes.c_uint64, P]
    L.o_splitmix_next.restype = ctypes.c_uint64
    L.o_splitmix_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    L.ontt_init.argtypes = [ctypes.POINTER(Ntt), ctypes.c_int]
    L.ontt_free.argtypes = [ctypes.POINTER(Ntt)]
    L.ontt_twiddle.restype = F128
    L.ontt_twiddle.argtypes = [ctypes.POINTER(Ntt), ctypes.c_int, SZ]
    L.ontt_forward_interleaved.argtypes = [ctypes.POINTER(Ntt), P, SZ, SZ, ctypes.c_int]
    L.ontt_inverse.argtypes = [ctypes.POINTER(Ntt), P, SZ]
    L.omerkle_tree.argtypes = [P, SZ, SZ, P]
    L.omerkle_multi_proof.restype = SZ
    L.omerkle_multi_proof.argtypes = [P, SZ, P, SZ, P, SZ]
    L.omerkle_verify_multi.argtypes = [P, SZ, P, P, SZ, P, SZ]
    L.o_pcs_commit.argtypes = [ctypes.POINTER(PcsParams), P, P, P, P]
    L.o_prove.argtypes = [P, SZ, P, SZ, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)), ctypes.POINTER(SZ), P]
    L.o_verify.argtypes = [P, SZ, SZ, P, SZ, P]
    L.o_statement_digest.argtypes = [ctypes.c_int, P]
    L.o_block_witness.argtypes = [P, P, P, P]
    L.o_witness_packed.argtypes = [P, SZ, ctypes.c_int, P, P, P]
    L.o_r1cs_matrices.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]
    L.o_apply_block.argtypes = [P, P, P]
    L.o_ext_matrix.argtypes = [P]
    L.o_lagrange_s.argtypes = [F128, P]
    L.o_lagrange_lambda.argtypes = [F128, P]
    L.free.argtypes = [P]


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ helpers
def f(x):
    return F128(x & 0xFFFFFFFFFFFFFFFF, x >> 64)


def i(v):
    return v.lo | (v.hi << 64)


def mul(a, b):
    return i(lib().of128_mul(f(a), f(b)))


def f128_array_to_ints(arr):
    a = arr.reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in a]


def ints_to_f128_array(vals):
    out = np.zeros((len(vals), 2), dtype=np.uint64)
    for k, v in enumerate(vals):
        out[k, 0] = v & 0xFFFFFFFFFFFFFFFF
        out[k, 1] = v >> 64
    return out.reshape(-1)


COMPRESSION_DTYPE = np.dtype([("cv", "<u4", 8), ("m", "<u4", 16), ("counter", "<u8"),
                              ("block_len", "<u4"), ("flags", "<u4")])


def generate_compressions(log2, seed):
    out = np.zeros(1 << log2, dtype=COMPRESSION_DTYPE)
    lib().o_generate_compressions(log2, seed, ptr(out))
    return out


def blake3(data):
    b = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().ob3_hash(ptr(b), len(data), ptr(out))
    return out.tobytes()


def leaf_cv(data):
    b = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().ob3_leaf_cv(ptr(b), len(data), ptr(out))
    return out.tobytes()


class NttHandle:
    def __init__(self, dim):
        self.t = Ntt()
        assert lib().ontt_init(ctypes.byref(self.t), dim) == 0

    def twiddle(self, layer, block):
        return i(lib().ontt_twiddle(ctypes.byref(self.t), layer, block))

    def forward_interleaved(self, data_u64, num_ntts, start_layer=0):
        n_total = data_u64.size // 2
        lib().ontt_forward_interleaved(ctypes.byref(self.t), ptr(data_u64), n_total, num_ntts, start_layer)

    def inverse(self, data_u64):
        lib().ontt_inverse(ctypes.byref(self.t), ptr(data_u64), data_u64.size // 2)

    def __del__(self):
        try:
            lib().ontt_free(ctypes.byref(self.t))
        except Exception:
            pass


def merkle_tree(data_u8, num_leaves):
    tree = np.zeros((2 * num_leaves - 1) * 32, np.uint8)
    lib().omerkle_tree(ptr(data_u8), data_u8.size, num_leaves, ptr(tree))
    return tree


def pcs_commit(m, log_inv_rate, log_batch_size, z_packed_u64, want_codeword=False):
    p = PcsParams(m, log_inv_rate, log_batch_size, 0, 1)
    log_dim = m - 7 - log_batch_size
    npos = 1 << (log_dim + log_inv_rate)
    root = np.zeros(32, np.uint8)
    cw = np.zeros(npos * (1 << log_batch_size) * 2, np.uint64) if want_codeword else None
    tree = np.zeros((2 * npos - 1) * 32, np.uint8) if want_codeword else None
    rc = lib().o_pcs_commit(ctypes.byref(p), ptr(z_packed_u64), ptr(cw) if cw is not None else None,
                            ptr(tree) if tree is not None else None, ptr(root))
    assert rc == 0
    return root.tobytes(), cw, tree


DOMAIN = b"flock-bench-v0"


def prove(blocks, domain=DOMAIN):
    """o_prove over a COMPRESSION_DTYPE array; returns the proof bytes."""
    L = lib()
    out = ctypes.POINTER(ctypes.c_uint8)()
    n = SZ()
    d = np.frombuffer(domain, np.uint8).copy()
    rc = L.o_prove(ptr(blocks), blocks.size, ptr(d), len(domain), ctypes.byref(out), ctypes.byref(n), None)
    assert rc == 0, rc
    data = ctypes.string_at(out, n.value)
    libc = ctypes.CDLL(None)
    libc.free(out)
    return data


def verify(proof, n_blocks, blocks=None, domain=DOMAIN):
    b = np.frombuffer(proof, np.uint8).copy()
    d = np.frombuffer(domain, np.uint8).copy()
    return lib().o_verify(ptr(b), b.size, n_blocks, ptr(d), len(domain), ptr(blocks) if blocks is not None else None)


def statement_digest(m):
    out = np.zeros(32, np.uint8)
    lib().o_statement_digest(m, ptr(out))
    return out.tobytes()


def block_witness(comp):
    z = np.zeros(256, np.uint64); a = np.zeros(256, np.uint64); b = np.zeros(256, np.uint64)
    c = np.array([comp], dtype=COMPRESSION_DTYPE)
    lib().o_block_witness(ptr(c), ptr(z), ptr(a), ptr(b))
    return z, a, b
