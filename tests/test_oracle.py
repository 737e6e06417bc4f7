"""Pins the oracle (oracle/cmref.c) before it is trusted as the checker.

Each check is independent of the oracle's own code: closed forms computed here in plain
Python, field laws, the reference decoder built from its own C sources (oracle/_ref),
its driver's errorstream files, and the golden fixtures made from them.  CPU only.
"""
import os
import tempfile

import numpy as np
import pytest

import _oracle as orc

PIDS = list(orc.PARAMS)
needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built (reference tree absent)")


# ---- independent textbook GF(2^m) in Python -------------------------------------------------

POLY = {12: (1 << 12) | 0x9, 13: (1 << 13) | 0x1B}  # x^12+x^3+1, x^13+x^4+x^3+x+1


def tb_mul(m, a, b):
    """Schoolbook carry-less product, then long division by the modulus."""
    prod = 0
    for i in range(m):
        if (b >> i) & 1:
            prod ^= a << i
    for k in range(2 * m - 2, m - 1, -1):
        if (prod >> k) & 1:
            prod ^= POLY[m] << (k - m)
    return prod


def tb_pow(m, a, e):
    r = 1
    for _ in range(e):
        r = tb_mul(m, r, a)
    return r


def tb_inv(m, a):
    # brute force over the (small) field for a spot check
    for x in range(1, 1 << m):
        if tb_mul(m, a, x) == 1:
            return x
    return 0


def np_clmul_mod(m, a, b):
    """Vectorised carry-less multiply + reduction (numpy, chunk-safe)."""
    a = a.astype(np.uint32)
    b = b.astype(np.uint32)
    prod = np.zeros_like(a)
    for i in range(m):
        prod ^= np.where((b >> i) & 1, a << i, 0).astype(np.uint32)
    for k in range(2 * m - 2, m - 1, -1):
        hit = (prod >> k) & 1
        prod ^= hit * np.uint32(POLY[m] << (k - m))
    return prod.astype(np.uint16)


@pytest.mark.parametrize("m", [12, 13])
def test_gf_mul_exhaustive(m):
    q = 1 << m
    b = np.arange(q, dtype=np.uint16)
    step = 256 if m == 13 else 512
    for a0 in range(0, q, step):
        a = np.repeat(np.arange(a0, a0 + step, dtype=np.uint16), q)
        bb = np.tile(b, step)
        assert np.array_equal(orc.gf_mul_batch(m, a, bb), np_clmul_mod(m, a, bb)), a0


@pytest.mark.parametrize("m", [12, 13])
def test_gf_field_laws(m):
    rng = np.random.default_rng(m)
    a, b, c = (rng.integers(0, 1 << m, 20000, dtype=np.uint16) for _ in range(3))
    mul = lambda x, y: orc.gf_mul_batch(m, x, y)
    assert np.array_equal(mul(a, b), mul(b, a))
    assert np.array_equal(mul(a, mul(b, c)), mul(mul(a, b), c))
    assert np.array_equal(mul(a, b ^ c), mul(a, b) ^ mul(a, c))
    assert np.array_equal(mul(a, np.ones_like(a)), a)
    # the multiplicative group has order 2^m - 1
    x = np.full(1, 2, np.uint16)
    assert tb_pow(m, 2, (1 << m) - 1) == 1
    assert int(orc.gf_mul_batch(m, x, x)[0]) == tb_mul(m, 2, 2)


@pytest.mark.parametrize("m", [12, 13])
def test_gf_inverse(m):
    a = np.arange(1 << m, dtype=np.uint16)
    inv = orc.gf_inv_batch(m, a)
    assert inv[0] == 0
    assert np.all(orc.gf_mul_batch(m, a[1:], inv[1:]) == 1)
    for v in (1, 2, 3, 0x1FF, (1 << m) - 1):
        assert int(inv[v]) == tb_inv(m, v)


@pytest.mark.parametrize("m", [12, 13])
def test_bitrev_involution(m):
    for a in range(0, 1 << m, 37):
        r = orc.bitrev(m, a)
        assert r < (1 << m)
        assert orc.bitrev(m, r) == a
        assert r == int(format(a, f"0{m}b")[::-1], 2)


@needs_ref
@pytest.mark.parametrize("pid", [348864, 8192128])
def test_gf_vs_reference_build(pid):
    ref = orc.Ref(pid)
    m = ref.m
    q = 1 << m
    b = np.arange(q, dtype=np.uint16)
    for a0 in range(0, q, 1024):
        a = np.repeat(np.arange(a0, a0 + 1024, dtype=np.uint16), q)
        bb = np.tile(b, 1024)
        assert np.array_equal(orc.gf_mul_batch(m, a, bb), ref.gf_mul_batch(a, bb))
    a = np.arange(q, dtype=np.uint16)
    assert np.array_equal(orc.gf_inv_batch(m, a), ref.gf_inv_batch(a))
    num = np.random.default_rng(1).integers(0, q, q, dtype=np.uint16)
    want = orc.gf_mul_batch(m, orc.gf_inv_batch(m, a), num)
    assert np.array_equal(ref.gf_frac_batch(a, num), want)  # gf_frac(den, num) = num/den
    assert np.array_equal(ref.bitrev_batch(a), np.array([orc.bitrev(m, int(v)) for v in a], np.uint16))


# ---- Benes network / support ---------------------------------------------------------------

@needs_ref
@pytest.mark.parametrize("pid", [348864, 460896])
def test_apply_benes_vs_reference_build(pid):
    ref = orc.Ref(pid)
    rng = np.random.default_rng(pid)
    for rev in (0, 1):
        for _ in range(4):
            cond = rng.integers(0, 256, ref.cond_bytes, dtype=np.uint8)
            r = rng.integers(0, 256, (1 << ref.m) // 8, dtype=np.uint8)
            assert np.array_equal(orc.apply_benes(ref.m, r, cond, rev), ref.apply_benes(r, cond, rev))


@pytest.mark.parametrize("m", [12, 13])
def test_apply_benes_inverse(m):
    rng = np.random.default_rng(m)
    cond = rng.integers(0, 256, (2 * m - 1) * (1 << (m - 4)), dtype=np.uint8)
    r = rng.integers(0, 256, (1 << m) // 8, dtype=np.uint8)
    fwd = orc.apply_benes(m, r, cond, 0)
    assert np.array_equal(orc.apply_benes(m, fwd, cond, 1), r)
    # a permutation network preserves the number of set bits
    assert np.unpackbits(fwd).sum() == np.unpackbits(r).sum()


@pytest.mark.parametrize("pid", PIDS)
def test_support_shipped_keys_golden(pid):
    """L from the reference's own support_gen (golden) and distinct elements."""
    gold = np.load(os.path.join(orc.GOLDEN, f"golden_{pid}.npz"))
    sk = orc.shipped_sk(pid)
    g, L, _ = orc.key_expand(pid, sk)
    assert np.array_equal(L, gold["L"])
    assert np.array_equal(g, gold["g"])
    assert len(np.unique(L)) == L.size


@pytest.mark.parametrize("pid", [348864, 6960119])
def test_support_roundtrip_controlbits(pid):
    """support_gen(controlbits(pi))[i] == bitrev(pi[i]) (the keypair convention,
    kem/common/core/pk_gen.c:69-70), with condition bits from the library's router."""
    import cm_decrypt__amd as cm
    m = orc.PARAMS[pid][0]
    pi = np.random.default_rng(pid).permutation(1 << m).astype(np.int16)
    cond = cm.controlbits(m, pi)
    L = orc.support_gen(pid, cond)
    assert np.array_equal(L, np.array([orc.bitrev(m, int(v)) for v in pi[:L.size]], np.uint16))


# ---- key table, syndrome, BM, root -----------------------------------------------------------

@pytest.mark.parametrize("pid", [348864, 6960119])
def test_table_closed_form(pid):
    m, n, t, ctb, sb, skb = orc.PARAMS[pid]
    g, L, T = orc.key_expand(pid, orc.shipped_sk(pid))
    assert T.shape == (sb, 2 * t)
    for i in (0, 1, 77, sb - 1):
        gl = 0
        for c in reversed(g.tolist()):
            gl = tb_mul(m, gl, int(L[i])) ^ c
        inv = tb_inv(m, tb_mul(m, gl, gl))
        for j in (0, 1, 5, 2 * t - 1):
            assert int(T[i, j]) == tb_mul(m, tb_pow(m, int(L[i]), j), inv), (i, j)


@needs_ref
@pytest.mark.parametrize("pid", PIDS)
def test_key_and_stages_vs_reference_build(pid):
    ref = orc.Ref(pid)
    sk = orc.shipped_sk(pid)
    ref.set_key(sk)
    g1, L1, T1 = ref.key()
    g, L, T = orc.key_expand(pid, sk)
    assert np.array_equal(g, g1) and np.array_equal(L, L1) and np.array_equal(T, T1)
    rng = np.random.default_rng(pid)
    ct = rng.integers(0, 256, (4, ref.ct_bytes), dtype=np.uint8)
    for c in ct:
        s = orc.synd(pid, T, c)
        assert np.array_equal(s, ref.synd(c))
        loc = orc.bm(pid, s)
        assert np.array_equal(loc, ref.bm(s))          # non-codeword input: only the reference pins it
        assert np.array_equal(orc.root(pid, loc, L), ref.root(loc))
    s_rand = rng.integers(0, 1 << ref.m, (3, 2 * ref.t), dtype=np.uint16)
    for s in s_rand:
        assert np.array_equal(orc.bm(pid, s), ref.bm(s))


@pytest.mark.parametrize("pid", [348864, 8192128])
def test_synd_linear_and_unit_rows(pid):
    m, n, t, ctb, sb, skb = orc.PARAMS[pid]
    g, L, T = orc.key_expand(pid, orc.shipped_sk(pid))
    rng = np.random.default_rng(2)
    a, b = rng.integers(0, 256, (2, ctb), dtype=np.uint8)
    assert np.array_equal(orc.synd(pid, T, a ^ b), orc.synd(pid, T, a) ^ orc.synd(pid, T, b))
    for i in (0, 9, sb - 1):
        unit = np.zeros(ctb, np.uint8)
        unit[i >> 3] = 1 << (i & 7)
        assert np.array_equal(orc.synd(pid, T, unit), T[i])


def poly_mul(m, a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] ^= tb_mul(m, x, y)
    return out


@pytest.mark.parametrize("pid,w", [(348864, 64), (348864, 9), (6960119, 119), (6960119, 30)])
def test_bm_closed_form_true_errors(pid, w):
    """For syndromes of a weight-w error at support points a_i, BM returns
    C(x) = prod (1 - a_i x) (coefficients above w zero), reversed (bm.c:66-67)."""
    m, n, t, ctb, sb, skb = orc.PARAMS[pid]
    g, L, _ = orc.key_expand(pid, orc.shipped_sk(pid))
    rng = np.random.default_rng(w)
    pos = rng.choice(n, size=w, replace=False)
    s = np.zeros(2 * t, np.uint16)
    for i in pos:
        a = int(L[i])
        gl = 0
        for c in reversed(g.tolist()):
            gl = tb_mul(m, gl, a) ^ c
        inv = tb_inv(m, tb_mul(m, gl, gl))
        p = inv
        for j in range(2 * t):
            s[j] ^= p
            p = tb_mul(m, p, a)
    C = [1]
    for i in pos:
        C = poly_mul(m, C, [1, int(L[i])])
    C = C + [0] * (t + 1 - len(C))
    assert np.array_equal(orc.bm(pid, s), np.array(C[::-1], np.uint16))


@pytest.mark.parametrize("pid", PIDS)
def test_decrypt_golden(pid):
    """Golden fixtures: shipped key, 5 generated ciphertexts, the reference driver's errorstream."""
    gold = np.load(os.path.join(orc.GOLDEN, f"golden_{pid}.npz"))
    e = orc.decrypt(pid, orc.shipped_sk(pid), gold["ct"])
    assert np.array_equal(e, gold["e"])
    text = b"".join(orc.errorstream_line(pid, x) for x in e)
    assert text == gold["stream"].tobytes()


def test_decrypt_golden_c1():
    gold = np.load(os.path.join(orc.GOLDEN, "golden_c1_348864.npz"))
    e = orc.decrypt(348864, gold["sk"], gold["ct"])
    assert np.array_equal(e, gold["e"])
    assert np.unpackbits(e, axis=1).sum(1).tolist() == [64] * 5
    assert b"".join(orc.errorstream_line(348864, x) for x in e) == gold["stream"].tobytes()


@needs_ref
@pytest.mark.parametrize("pid", [348864, 6960119, 8192128])
def test_reference_driver_errorstream(pid):
    """The reference driver's file output equals the oracle's lines byte for byte, also for
    ciphertexts of weight < t and random (non-codeword) ciphertexts."""
    import cm_decrypt__amd as cm
    sk = cm.gen_keypair(pid, 55)
    ct, _ = cm.gen_encrypt(pid, sk, 3, 3)
    ct_low, _ = cm.gen_encrypt(pid, sk, 4, 2, weight=orc.PARAMS[pid][2] - 7)
    ct_rand = np.random.default_rng(1).integers(0, 256, (2, orc.PARAMS[pid][3]), dtype=np.uint8)
    cts = np.concatenate([ct, ct_low, ct_rand])
    with tempfile.TemporaryDirectory() as d:
        rc, _, stream = orc.run_ref_driver(pid, sk, cts, d)
    assert rc == 0
    e = orc.decrypt(pid, sk, cts)
    assert stream == b"".join(orc.errorstream_line(pid, x) for x in e)


@needs_ref
def test_reference_driver_bad_katnum():
    """KATNUM <= 0 or non-numeric: usage and KAT_CRYPTO_FAILURE (decrypt.c:26-41, 121-123)."""
    import subprocess
    exe = os.path.join(orc.REF_DIR, "bin", "decrypt_6960119")
    for arg in ("0", "-3", "abc", ""):
        r = subprocess.run([exe, arg], capture_output=True)
        assert r.returncode == (-4) & 0xFF


@pytest.mark.parametrize("pid,K", [(348864, 63), (8192128, 63), (8192128, 127)])
def test_late_first_syndrome_vs_reference_build(pid, K):
    """Ciphertexts whose syndrome is zero below index K (K = 63, 127): the first nonzero
    discrepancy comes when B = x^(K+1), a coefficient on a 64-lane block boundary of the
    GPU layouts.  Oracle and reference build agree on them, stage by stage."""
    if not orc.ref_available(pid):
        pytest.skip("reference build not present")
    from _gf2 import late_syndrome_cts
    sk = orc.shipped_sk(pid)
    ct = late_syndrome_cts(pid, sk, K, 3, seed=K)
    ref = orc.Ref(pid)
    ref.set_key(sk)
    _, _, T = orc.key_expand(pid, sk)
    for c in ct:
        s = orc.synd(pid, T, c)
        assert not np.any(s[:K]) and s[K] != 0
        assert np.array_equal(s, ref.synd(c))
        assert np.array_equal(orc.bm(pid, s), ref.bm(s))
    assert np.array_equal(orc.decrypt(pid, sk, ct), ref.decrypt(ct))


# ---- decapsulation pieces -------------------------------------------------------------------

def test_shake256_known_answer():
    """SHAKE256 of the empty message, 256 output bits (FIPS 202 example value)."""
    assert orc.shake256_32(b"").hex() == "46b9dd2b0ba88d13233b3feb743eeb243fcd52ea62b81b82b50c27646ed5762f"


@pytest.mark.parametrize("pid", PIDS)
def test_synd_full_is_reference_synd_on_ciphertext_bits(pid):
    """On words inside the first sb positions the full-support syndrome equals the reference's
    table-driven synd (util.c:102-114)."""
    m, n, t, ctb, sb, _ = orc.PARAMS[pid]
    g, L, T = orc.key_expand(pid, orc.shipped_sk(pid))
    rng = np.random.default_rng(pid)
    for _ in range(3):
        ct = rng.integers(0, 256, ctb, dtype=np.uint8)
        r = np.zeros(n // 8, np.uint8)
        r[:ctb] = ct
        assert np.array_equal(orc.synd_full(pid, g, L, r), orc.synd(pid, T, ct))


@pytest.mark.parametrize("pid", PIDS)
def test_synd_full_codeword_relation(pid):
    """c = H e for the golden ciphertexts: s(e) over all n positions equals s(c); a flipped bit
    of e outside the first sb positions breaks the equality."""
    m, n, t, ctb, sb, _ = orc.PARAMS[pid]
    gold = np.load(os.path.join(orc.GOLDEN, f"golden_{pid}.npz"))
    g, L, T = orc.key_expand(pid, orc.shipped_sk(pid))
    for c, e in zip(gold["ct"], gold["e"]):
        assert np.array_equal(orc.synd_full(pid, g, L, e), orc.synd(pid, T, c))
    e2 = gold["e"][0].copy()
    e2[-1] ^= 0x80
    assert not np.array_equal(orc.synd_full(pid, g, L, e2), orc.synd(pid, T, gold["ct"][0]))


@pytest.mark.parametrize("pid", PIDS)
def test_kem_dec_oracle(pid):
    """Valid ciphertexts give the reference encapsulation's key SHAKE256(1 || e || c)
    (kem/common/core/operations.c:15-34); random ciphertexts, a flipped ciphertext bit and
    errors of weight t-1 are rejected with SHAKE256(0 || s || c)."""
    m, n, t, ctb, sb, _ = orc.PARAMS[pid]
    sk = orc.shipped_sk(pid)
    gold = np.load(os.path.join(orc.GOLDEN, f"golden_{pid}.npz"))
    keys, ok = orc.kem_dec(pid, sk, gold["ct"][:3])
    assert ok.all()
    for c, e, k in zip(gold["ct"], gold["e"], keys):
        assert k.tobytes() == orc.shake256_32(b"\x01" + e.tobytes() + c.tobytes())
    # a ciphertext bit that is also an error position: decodes to weight t-1
    c0, e0 = gold["ct"][0].copy(), gold["e"][0]
    i = int(np.flatnonzero(np.unpackbits(e0[:ctb], bitorder="little"))[0])
    low = c0.copy()
    low[i >> 3] ^= 1 << (i & 7)
    flip = c0.copy()
    flip[0] ^= 0x01 if not (e0[0] & 1) else 0x02
    rnd = np.random.default_rng(pid).integers(0, 256, (2, ctb), dtype=np.uint8)
    pad = (m * t) % 8
    if pad:
        rnd[:, -1] &= (1 << pad) - 1  # padding bits clear: rejected by the decoder's check
    bad = np.vstack([low, flip, rnd])
    keys, ok = orc.kem_dec(pid, sk, bad)
    assert not ok.any()
    s_rej = sk[40 + 2 * t + orc.COND_BYTES[pid]:][:n // 8].tobytes()
    for c, k in zip(bad, keys):
        assert k.tobytes() == orc.shake256_32(b"\x00" + s_rej + c.tobytes())
    if pad:
        # a valid ciphertext with a padding bit set: all-zero key (the specification's rule)
        c1 = gold["ct"][0].copy()
        c1[-1] |= np.uint8(1 << pad)
        keys, ok = orc.kem_dec(pid, sk, c1[None, :])
        assert not ok[0] and not keys.any()
