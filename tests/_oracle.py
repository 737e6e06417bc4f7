"""ctypes access to the checker: oracle/liboracle.so (C restatement of the reference
decoder) and oracle/_ref/libref_<p>.so (the reference CPU decoder built from its own
sources, present where it was built).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
GOLDEN = os.path.join(ROOT, "tests", "golden")

PARAMS = {  # id: (m, n, t, ct_bytes, sb, sk_bytes)
    348864: (12, 3488, 64, 96, 768, 6492),
    460896: (13, 4608, 96, 156, 1248, 13608),
    6688128: (13, 6688, 128, 208, 1664, 13932),
    6960119: (13, 6960, 119, 194, 1552, 13948),
    8192128: (13, 8192, 128, 208, 1664, 14120),
}

_orc = None


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _orc = ctypes.CDLL(ORACLE_SO)
        _orc.orc_timed_decrypt.restype = ctypes.c_double
        for f in ("orc_gf_mul", "orc_gf_inv", "orc_gf_frac", "orc_bitrev"):
            getattr(_orc, f).restype = ctypes.c_uint16
        _orc.orc_errorstream_line.restype = ctypes.c_long
    return _orc


def gf_mul_batch(m, a, b):
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    out = np.zeros_like(a)
    lib().orc_gf_mul_batch(m, _p(a), _p(b), _p(out), ctypes.c_size_t(a.size))
    return out


def gf_inv_batch(m, a):
    a = np.ascontiguousarray(a, np.uint16)
    out = np.zeros_like(a)
    lib().orc_gf_inv_batch(m, _p(a), _p(out), ctypes.c_size_t(a.size))
    return out


def bitrev(m, a):
    return int(lib().orc_bitrev(m, ctypes.c_uint16(a)))


def apply_benes(m, r, cond, rev=0):
    r = np.array(r, dtype=np.uint8, copy=True)
    cond = np.ascontiguousarray(cond, np.uint8)
    assert lib().orc_apply_benes(m, _p(r), _p(cond), rev) == 0
    return r


def support_gen(pid, cond):
    n = PARAMS[pid][1]
    L = np.zeros(n, np.uint16)
    assert lib().orc_support_gen(pid, _p(np.ascontiguousarray(cond, np.uint8)), _p(L)) == 0
    return L


def key_expand(pid, sk):
    m, n, t, ctb, sb, skb = PARAMS[pid]
    g = np.zeros(t + 1, np.uint16)
    L = np.zeros(n, np.uint16)
    T = np.zeros((sb, 2 * t), np.uint16)
    assert lib().orc_key_expand(pid, _p(np.ascontiguousarray(sk, np.uint8)), _p(g), _p(L), _p(T)) == 0
    return g, L, T


def synd(pid, table, ct):
    t = PARAMS[pid][2]
    s = np.zeros(2 * t, np.uint16)
    assert lib().orc_synd(pid, _p(np.ascontiguousarray(table, np.uint16)), _p(np.ascontiguousarray(ct, np.uint8)),
                          _p(s)) == 0
    return s


def synd_full(pid, g, L, r):
    """s_j(r) = sum_{i<n} r_i L_i^j / g(L_i)^2, j < 2t (all n support positions)."""
    t = PARAMS[pid][2]
    s = np.zeros(2 * t, np.uint16)
    assert lib().orc_synd_full(pid, _p(np.ascontiguousarray(g, np.uint16)), _p(np.ascontiguousarray(L, np.uint16)),
                               _p(np.ascontiguousarray(r, np.uint8)), _p(s)) == 0
    return s


COND_BYTES = {348864: 5888, 460896: 12800, 6688128: 12800, 6960119: 12800, 8192128: 12800}


def shake256_32(data):
    import hashlib
    return hashlib.shake_256(bytes(data)).digest(32)


def kem_dec(pid, sk, ct):
    """crypto_kem_dec of the Classic McEliece specification for ciphertext rows under one key:
    (session keys count x 32, accepted flags).  e from the oracle decoder is accepted when its
    weight is t and s(e) over all n positions equals the ciphertext's syndrome; the key is
    SHAKE256(1 || e || c) then (the reference encapsulation's formula,
    kem/common/core/operations.c:15-34), else SHAKE256(0 || s || c) with the secret key's s.
    For 6960119 a ciphertext with nonzero padding bits (above bit mt of its last byte) gets an
    all-zero key, the specification's rule for that set; the reference has no decapsulation,
    so this branch is pinned by the specification alone."""
    m, n, t, ctb, sb, skb = PARAMS[pid]
    sk = np.ascontiguousarray(sk, np.uint8)
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, ctb)
    g, L, T = key_expand(pid, sk)
    s_rej = sk[40 + 2 * t + COND_BYTES[pid]:][:n // 8]
    E = decrypt(pid, sk, ct)
    keys = np.zeros((ct.shape[0], 32), np.uint8)
    ok = np.zeros(ct.shape[0], bool)
    for i in range(ct.shape[0]):
        e = E[i]
        w = int(np.unpackbits(e).sum())
        good = w == t and np.array_equal(synd(pid, T, ct[i]), synd_full(pid, g, L, e))
        pre = (bytes([1]) + e.tobytes() if good else bytes([0]) + s_rej.tobytes()) + ct[i].tobytes()
        keys[i] = np.frombuffer(shake256_32(pre), np.uint8)
        ok[i] = good
        if (m * t) % 8 and int(ct[i, ctb - 1]) >> ((m * t) % 8):
            # nonzero padding bits (6960119): the specification's decapsulation returns an
            # all-zero key for such a ciphertext
            keys[i] = 0
            ok[i] = False
    return keys, ok


def bm(pid, s):
    t = PARAMS[pid][2]
    out = np.zeros(t + 1, np.uint16)
    assert lib().orc_bm(pid, _p(np.ascontiguousarray(s, np.uint16)), _p(out)) == 0
    return out


def root(pid, loc, L):
    n = PARAMS[pid][1]
    img = np.zeros(n, np.uint16)
    assert lib().orc_root(pid, _p(np.ascontiguousarray(loc, np.uint16)), _p(np.ascontiguousarray(L, np.uint16)),
                          _p(img)) == 0
    return img


def error_vector(pid, images):
    n = PARAMS[pid][1]
    e = np.zeros(n // 8, np.uint8)
    assert lib().orc_error_vector(pid, _p(np.ascontiguousarray(images, np.uint16)), _p(e)) == 0
    return e


def decrypt(pid, sk, ct):
    """Error vectors (count x n/8) for ciphertext rows under one secret key."""
    m, n, t, ctb, sb, skb = PARAMS[pid]
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, ctb)
    e = np.zeros((ct.shape[0], n // 8), np.uint8)
    assert lib().orc_decrypt(pid, _p(np.ascontiguousarray(sk, np.uint8)), _p(ct), ctypes.c_size_t(ctb),
                             ctypes.c_size_t(ct.shape[0]), _p(e)) == 0
    return e


def timed_decrypt(pid, sk, ct):
    m, n, t, ctb, sb, skb = PARAMS[pid]
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, ctb)
    e = np.zeros((ct.shape[0], n // 8), np.uint8)
    stage = np.zeros(3, np.float64)
    wall = lib().orc_timed_decrypt(pid, _p(np.ascontiguousarray(sk, np.uint8)), _p(ct), ctypes.c_size_t(ctb),
                                   ctypes.c_size_t(ct.shape[0]), _p(e), _p(stage))
    return wall, stage, e


def errorstream_line(pid, e):
    buf = ctypes.create_string_buffer(1 << 16)
    k = lib().orc_errorstream_line(pid, _p(np.ascontiguousarray(e, np.uint8)), buf, ctypes.c_size_t(1 << 16))
    assert k >= 0
    return buf.raw[:k]


class Ref:
    """The reference CPU decoder for one parameter set (oracle/_ref/libref_<p>.so)."""

    def __init__(self, pid):
        path = os.path.join(REF_DIR, f"libref_{pid}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.pid = pid
        self.lib = ctypes.CDLL(path)
        self.lib.ref_timed_decrypt.restype = ctypes.c_double
        info = (ctypes.c_int * 7)()
        self.lib.ref_params(info)
        self.m, self.n, self.t, self.ct_bytes, self.sb, self.sk_bytes, self.cond_bytes = list(info)

    def gf_mul_batch(self, a, b):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        out = np.zeros_like(a)
        self.lib.ref_gf_mul_batch(_p(a), _p(b), _p(out), ctypes.c_long(a.size))
        return out

    def gf_inv_batch(self, a):
        a = np.ascontiguousarray(a, np.uint16)
        out = np.zeros_like(a)
        self.lib.ref_gf_inv_batch(_p(a), _p(out), ctypes.c_long(a.size))
        return out

    def gf_frac_batch(self, den, num):
        den = np.ascontiguousarray(den, np.uint16)
        num = np.ascontiguousarray(num, np.uint16)
        out = np.zeros_like(den)
        self.lib.ref_gf_frac_batch(_p(den), _p(num), _p(out), ctypes.c_long(den.size))
        return out

    def bitrev_batch(self, a):
        a = np.ascontiguousarray(a, np.uint16)
        out = np.zeros_like(a)
        self.lib.ref_bitrev_batch(_p(a), _p(out), ctypes.c_long(a.size))
        return out

    def apply_benes(self, r, cond, rev=0):
        r = np.array(r, dtype=np.uint8, copy=True)
        self.lib.ref_apply_benes(_p(r), _p(np.ascontiguousarray(cond, np.uint8)), rev)
        return r

    def support_gen(self, cond):
        L = np.zeros(self.n, np.uint16)
        self.lib.ref_support_gen(_p(np.ascontiguousarray(cond, np.uint8)), _p(L))
        return L

    def set_key(self, sk):
        self.lib.ref_set_key(_p(np.ascontiguousarray(sk, np.uint8)))

    def key(self):
        g = np.zeros(self.t + 1, np.uint16)
        L = np.zeros(self.n, np.uint16)
        T = np.zeros((self.sb, 2 * self.t), np.uint16)
        self.lib.ref_get_key(_p(g), _p(L), _p(T))
        return g, L, T

    def synd(self, ct):
        s = np.zeros(2 * self.t, np.uint16)
        self.lib.ref_synd(_p(np.ascontiguousarray(ct, np.uint8)), _p(s))
        return s

    def bm(self, s):
        out = np.zeros(self.t + 1, np.uint16)
        self.lib.ref_bm(_p(np.ascontiguousarray(s, np.uint16)), _p(out))
        return out

    def root(self, f):
        img = np.zeros(self.n, np.uint16)
        self.lib.ref_root(_p(np.ascontiguousarray(f, np.uint16)), _p(img))
        return img

    def decrypt(self, ct):
        ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, self.ct_bytes)
        e = np.zeros((ct.shape[0], self.n // 8), np.uint8)
        self.lib.ref_decrypt(_p(ct), ctypes.c_long(self.ct_bytes), ctypes.c_long(ct.shape[0]), _p(e))
        return e

    def timed_decrypt(self, ct):
        ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, self.ct_bytes)
        e = np.zeros((ct.shape[0], self.n // 8), np.uint8)
        wall = self.lib.ref_timed_decrypt(_p(ct), ctypes.c_long(self.ct_bytes), ctypes.c_long(ct.shape[0]), _p(e))
        return wall, e


def ref_available(pid=8192128):
    return os.path.exists(os.path.join(REF_DIR, f"libref_{pid}.so"))


def shipped_sk(pid):
    """Secret key shipped with the reference (kem/test_vectors/Cipher_Sk), kept as a fixture."""
    return np.fromfile(os.path.join(GOLDEN, f"sk_{pid}.bin"), dtype=np.uint8)


def run_ref_driver(pid, sk, ct, workdir, katnum=None):
    """Runs the reference CPU driver (oracle/_ref/bin/decrypt_<p>, built from the reference's
    sources) on ct/sk files laid out the way it expects (util.c:12, decrypt.c:18):
    <root>/kem/test_vectors/Cipher_Sk/{ct,sk}_<p>.bin read from cwd <root>/CPU/param/param<p>,
    output <root>/CPU/results/output/errorstream0_<p>.bin.  Returns (returncode, stdout, errorstream)."""
    import subprocess
    exe = os.path.join(REF_DIR, "bin", f"decrypt_{pid}")
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    vec = os.path.join(workdir, "kem", "test_vectors", "Cipher_Sk")
    cwd = os.path.join(workdir, "CPU", "param", f"param{pid}")
    out_dir = os.path.join(workdir, "CPU", "results", "output")
    for d in (vec, cwd, out_dir):
        os.makedirs(d, exist_ok=True)
    np.ascontiguousarray(sk, np.uint8).tofile(os.path.join(vec, f"sk_{pid}.bin"))
    ct = np.ascontiguousarray(ct, np.uint8)
    ct.tofile(os.path.join(vec, f"ct_{pid}.bin"))
    out = os.path.join(out_dir, f"errorstream0_{pid}.bin")
    if os.path.exists(out):
        os.remove(out)
    k = katnum if katnum is not None else ct.shape[0]
    r = subprocess.run([exe, str(k)], cwd=cwd, capture_output=True, timeout=600)
    data = open(out, "rb").read() if os.path.exists(out) else None
    return r.returncode, r.stdout.decode(errors="replace"), data
