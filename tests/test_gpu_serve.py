"""Serving loop through the C ABI: keys held against unload while requests are queued, a
failed launch failing only its own requests, and decapsulation requests (alone and mixed with
decryptions in the same launches) against the oracle's crypto_kem_dec."""
import numpy as np
import pytest

import cm_decrypt__amd as cm
import _oracle as orc

pytestmark = pytest.mark.gpu
PIDS = list(cm.PARAM_IDS)


def test_server_unload_refused_while_requests_queued(ctx):
    """Submission holds each request's key until it retires: unloading the key while 1024
    requests are queued returns CMD_ERR_KEY and the requests still decode correctly."""
    pid = 8192128
    sk = cm.gen_keypair(pid, 5151)
    kid = ctx.load_key(pid, sk)
    ct, e = cm.gen_encrypt(pid, sk, 77, 1024)
    # one launch slot of 256 items: the requests need four launches one after another
    srv = cm.Server(ctx, max_batch=256, max_inflight=1, min_batch=1, max_wait_us=0.0)
    try:
        rids = srv.submit_batch([kid] * len(ct), ct)
        with pytest.raises(cm.CmdError) as ei:
            ctx.unload_key(kid)
        assert ei.value.code == cm.CMD_ERR_KEY
        got = srv.wait_batch(rids, e.shape[1], 60e6)
    finally:
        srv.close()
    assert np.array_equal(got, e)
    ctx.unload_key(kid)  # every request has retired: the key can go
    srv = cm.Server(ctx, max_batch=64, max_inflight=1)
    try:
        with pytest.raises(cm.CmdError) as ei:
            srv.submit_batch([kid], ct[:1])
        assert ei.value.code == cm.CMD_ERR_KEY
    finally:
        srv.close()


def test_server_serves_after_failed_launch(ctx):
    """A launch whose decapsulation scratch does not fit the HBM budget fails its own requests
    with CMD_ERR_BUDGET; the server keeps serving, and later decryptions are exact."""
    pid = 348864
    sk = cm.gen_keypair(pid, 5252)
    kid = ctx.load_key(pid, sk)
    ct, e = cm.gen_encrypt(pid, sk, 78, 64)
    _, budget0 = ctx.memory()
    srv = cm.Server(ctx, max_batch=64, max_inflight=1, min_batch=1, max_wait_us=0.0)
    try:
        used, _ = ctx.memory()
        # a decapsulation launch needs at least 256 B of flags + 1360 B of SHAKE256 input
        ctx.set_budget(used + 1024)
        rk = srv.submit_batch([kid] * len(ct), ct, mode=cm.MODE_KEM)
        with pytest.raises(cm.CmdError) as ei:
            srv.wait_batch(rk, 32, 60e6)
        assert ei.value.code == cm.CMD_ERR_BUDGET
        print("failed launch reported; decrypting after it", flush=True)
        rd = srv.submit_batch([kid] * len(ct), ct)
        got = srv.wait_batch(rd, e.shape[1], 60e6)
        st = srv.stats()
    finally:
        ctx.set_budget(budget0)
        srv.close()
    assert np.array_equal(got, e)
    assert st["failed"] == len(ct) and st["completed"] == 2 * len(ct)
    ctx.unload_key(kid)


def kem_cases(pid, sk, seed):
    """Valid ciphertexts, random (rejected) ones and, for 6960119, valid ones with a padding
    bit set (all-zero key)."""
    p = cm.params(pid)
    ct, _ = cm.gen_encrypt(pid, sk, seed, 6)
    rnd = np.random.default_rng(seed).integers(0, 256, size=(2, p.ct_bytes), dtype=np.uint8)
    if (p.m * p.t) % 8:
        rnd[:, -1] &= (1 << ((p.m * p.t) % 8)) - 1
    rows = [ct, rnd]
    if (p.m * p.t) % 8:
        bad = ct[:2].copy()
        bad[:, -1] |= np.uint8(1 << 7)
        rows.append(bad)
    return np.concatenate(rows)


def test_server_decapsulations_alone_and_mixed(ctx):
    """Decapsulation requests through the server give the oracle's session keys, submitted
    alone and shuffled together with decryption requests of all five sets."""
    kem, dec = [], []
    kids = {}
    for pid in PIDS:
        sk = cm.gen_keypair(pid, 5300 + pid % 1000)
        kid = ctx.load_key(pid, sk)
        kids[pid] = kid
        cts = kem_cases(pid, sk, 60 + pid % 71)
        keys, _ = orc.kem_dec(pid, sk, cts)
        kem += [(kid, c, k) for c, k in zip(cts, keys)]
        dct, de = cm.gen_encrypt(pid, sk, 61 + pid % 73, 5)
        dec += [(kid, c, x) for c, x in zip(dct, de)]
    srv = cm.Server(ctx, max_batch=128, max_inflight=2, min_batch=1, max_wait_us=100.0)
    try:
        alone = []
        for kid, c, _ in kem:
            rid = srv.submit_batch([kid], c[None, :], mode=cm.MODE_KEM)
            assert srv.result_bytes(int(rid[0])) == 32
            alone.append(srv.wait(int(rid[0]), 32, 10e6))
        for (kid, c, want), got in zip(kem, alone):
            assert np.array_equal(got, want), f"key {kid}: decapsulation alone differs from the oracle"
        print(f"{len(kem)} decapsulations alone match the oracle", flush=True)
        tagged = [("k", i) for i in range(len(kem))] + [("d", i) for i in range(len(dec))]
        order = np.random.default_rng(11).permutation(len(tagged))
        ids = []
        for j in order:
            tag, i = tagged[j]
            kid, c, _ = (kem if tag == "k" else dec)[i]
            mode = cm.MODE_KEM if tag == "k" else cm.MODE_FUSED
            ids.append((tag, i, int(srv.submit_batch([kid], c[None, :], mode=mode)[0])))
        for tag, i, rid in ids:
            want = (kem if tag == "k" else dec)[i][2]
            got = srv.wait(rid, want.size, 10e6)
            assert np.array_equal(got, want), f"{tag} request {i} differs inside the mixed stream"
        st = srv.stats()
    finally:
        srv.close()
    assert st["failed"] == 0
    for kid in kids.values():
        ctx.unload_key(kid)
