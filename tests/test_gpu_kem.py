"""KEM decapsulation on the GPU (CMD_MODE_KEM, cmd_kem_dec, cmd_stage_synd_full) against the
oracle: the full-support syndrome of the transposed-FFT check kernel, the accept/reject
decision and the SHAKE256 session keys, for every parameter set, mixed batches and graph
replays."""
import numpy as np
import pytest
import torch

import cm_decrypt__amd as cm
import _oracle as orc

pytestmark = pytest.mark.gpu
PIDS = list(cm.PARAM_IDS)


@pytest.fixture(scope="module")
def kem_keys(ctx):
    out = {}
    for pid in PIDS:
        sk = cm.gen_keypair(pid, 7300 + pid % 1000)
        out[pid] = (sk, ctx.load_key(pid, sk))
    return out


@pytest.mark.parametrize("pid", PIDS)
def test_synd_full_vs_oracle(ctx, kem_keys, pid):
    """Dense random words (a full tile and a partial one), weight-t error vectors and zero."""
    sk, kid = kem_keys[pid]
    m, n, t = orc.PARAMS[pid][:3]
    g, L, _ = orc.key_expand(pid, sk)
    rng = np.random.default_rng(pid + 7)
    dense = rng.integers(0, 256, (37, n // 8), dtype=np.uint8)
    _, sparse = cm.gen_encrypt(pid, sk, 17, 5)
    r = np.vstack([dense, sparse, np.zeros((1, n // 8), np.uint8)])
    print(f"synd_full pid={pid} rows={len(r)}", flush=True)
    s = ctx.synd_full(kid, r)
    for i in range(len(r)):
        assert np.array_equal(s[i], orc.synd_full(pid, g, L, r[i])), f"row {i}"
    assert not s[-1].any()


def kem_cases(pid, sk, seed):
    """Ciphertexts with the expected acceptance: 34 valid, then weight t-1, a flipped bit of a
    valid ciphertext, and random bytes (all rejected)."""
    m, n, t, ctb = orc.PARAMS[pid][:4]
    ct, e = cm.gen_encrypt(pid, sk, seed, 34)
    low, _ = cm.gen_encrypt(pid, sk, seed + 1, 3, weight=t - 1)
    flip = ct[:3].copy()
    flip[:, 5] ^= 0x10
    rnd = np.random.default_rng(seed).integers(0, 256, (3, ctb), dtype=np.uint8)
    cts = np.vstack([ct, low, flip, rnd])
    return cts, e


@pytest.mark.parametrize("pid", PIDS)
def test_kem_dec_vs_oracle(ctx, kem_keys, pid):
    sk, kid = kem_keys[pid]
    cts, e = kem_cases(pid, sk, 40 + pid % 97)
    print(f"kem_dec pid={pid} items={len(cts)}", flush=True)
    got = ctx.kem_dec([kid] * len(cts), cts)
    want, ok = orc.kem_dec(pid, sk, cts)
    assert ok[:34].all() and not ok[34:].any()
    for i in range(34):  # the reference encapsulation's key for the encrypted e
        assert got[i].tobytes() == orc.shake256_32(b"\x01" + e[i].tobytes() + cts[i].tobytes())
    assert np.array_equal(got, want)


def test_kem_dec_mixed_batch_graph(ctx, kem_keys):
    """All five sets shuffled in one batch of padded 208-byte rows: cmd_kem_dec and a KEM
    batch captured as a hipGraph, replayed with other device work in between."""
    rows, kids, want = [], [], []
    for pid in PIDS:
        sk, kid = kem_keys[pid]
        cts, _ = kem_cases(pid, sk, 90 + pid % 53)
        keys, _ = orc.kem_dec(pid, sk, cts)
        for c, k in zip(cts, keys):
            r = np.zeros(208, np.uint8)
            r[:c.size] = c
            rows.append(r)
            kids.append(kid)
            want.append(k)
    order = np.random.default_rng(5).permutation(len(rows))
    ct = np.stack(rows)[order]
    kids = np.array(kids, np.int32)[order]
    want = np.stack(want)[order]
    print(f"kem mixed items={len(ct)}", flush=True)
    assert np.array_equal(ctx.kem_dec(kids, ct, ct_stride=208), want)
    d_ct = torch.from_numpy(ct).cuda()
    d_k = torch.zeros((len(ct), 48), dtype=torch.uint8, device="cuda")
    b = cm.Batch(ctx, kids, 208, 48, mode=cm.MODE_KEM, use_graph=True)
    try:
        for rep in range(3):
            d_k.zero_()
            junk = torch.randn(256, 256, device="cuda") @ torch.randn(256, 256, device="cuda")
            b.run(d_ct.data_ptr(), d_k.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            assert float(junk.abs().sum()) > 0
            out = d_k.cpu().numpy()
            assert np.array_equal(out[:, :32], want), f"replay {rep}"
            assert not out[:, 32:].any()
    finally:
        b.close()


def test_kem_dec_rejects_short_key_stride(ctx, kem_keys):
    sk, kid = kem_keys[348864]
    ct, _ = cm.gen_encrypt(348864, sk, 3, 2)
    with pytest.raises(cm.CmdError):
        b = cm.Batch(ctx, [kid, kid], 96, 16, mode=cm.MODE_KEM)
        b.close()


def test_kem_dec_nonzero_padding_6960119(ctx, kem_keys):
    """mceliece6960119 ciphertexts hold 1547 bits in 194 bytes: a valid ciphertext with any of
    its five padding bits set gets the all-zero key, while the same ciphertext with clean
    padding keeps the reference encapsulation's key.  The rule is the specification's for this
    set (the reference has no decapsulation)."""
    pid = 6960119
    sk, kid = kem_keys[pid]
    ct, e = cm.gen_encrypt(pid, sk, 123, 6)
    padded = ct.copy()
    for i in range(5):
        padded[i, -1] |= np.uint8(1 << (3 + i))  # bits 3..7 of the last byte are padding
    cts = np.vstack([ct, padded])
    print(f"kem padding items={len(cts)}", flush=True)
    got = ctx.kem_dec([kid] * len(cts), cts)
    want, ok = orc.kem_dec(pid, sk, cts)
    assert np.array_equal(got, want)
    assert ok[:6].all() and not ok[6:11].any() and ok[11]
    assert not got[6:11].any() and got[:6].any(axis=1).all()
    for i in range(6):
        assert got[i].tobytes() == orc.shake256_32(b"\x01" + e[i].tobytes() + ct[i].tobytes())
