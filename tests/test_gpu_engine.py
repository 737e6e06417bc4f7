"""GPU engine behaviour through the C ABI: graph replays, device-buffer entry points, the
serving loop (a request decodes the same alone and inside a mixed stream), key lifetime,
HBM budgets, schedule persistence and autotuning.  Results are checked against the oracle
or against the error vectors the generator encrypted."""
import json
import os
import tempfile

import numpy as np
import pytest
import torch

import cm_decrypt__amd as cm
import _oracle as orc

pytestmark = pytest.mark.gpu
PIDS = list(cm.PARAM_IDS)


@pytest.fixture(scope="module")
def keys(ctx):
    out = {}
    for pid in PIDS:
        sk = cm.gen_keypair(pid, 9100 + pid % 1000)
        out[pid] = (sk, ctx.load_key(pid, sk))
    return out


def mixed_rows(keys, per_key, seed):
    """Padded 208-byte ciphertext rows of every set, shuffled; expected 1024-byte error rows."""
    rows, kids, want = [], [], []
    for pid in PIDS:
        sk, kid = keys[pid]
        ct, e = cm.gen_encrypt(pid, sk, seed + pid % 89, per_key)
        for c, x in zip(ct, e):
            r = np.zeros(208, np.uint8)
            r[:c.size] = c
            w = np.zeros(1024, np.uint8)
            w[:x.size] = x
            rows.append(r)
            kids.append(kid)
            want.append(w)
    order = np.random.default_rng(seed).permutation(len(rows))
    return np.stack(rows)[order], np.array(kids, np.int32)[order], np.stack(want)[order]


@pytest.mark.parametrize("mode", [cm.MODE_PLAIN, cm.MODE_FUSED])
def test_batch_graph_replays(ctx, keys, mode):
    """A captured batch replays correctly several times with other device work and host
    synchronisation between replays, and re-captures when the buffers change."""
    ct, kids, want = mixed_rows(keys, 23, 61)
    d_ct = torch.from_numpy(ct).cuda()
    d_e = [torch.full((len(ct), 1024), 0xA5, dtype=torch.uint8, device="cuda") for _ in range(2)]
    for d in d_e:
        d.zero_()
    b = cm.Batch(ctx, kids, 208, 1024, mode=mode, use_graph=True)
    try:
        for rep in range(4):
            out = d_e[rep % 2]
            out.zero_()
            junk = torch.randn(512, 512, device="cuda") @ torch.randn(512, 512, device="cuda")
            b.run(d_ct.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            assert float(junk.abs().sum()) > 0
            assert np.array_equal(out.cpu().numpy(), want), f"replay {rep}"
    finally:
        b.close()


def test_decrypt_device_on_side_stream(ctx, keys):
    ct, kids, want = mixed_rows(keys, 11, 62)
    d_ct = torch.from_numpy(ct).cuda()
    d_e = torch.zeros((len(ct), 1024), dtype=torch.uint8, device="cuda")
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    for mode in (cm.MODE_PLAIN, cm.MODE_FUSED):
        d_e.zero_()
        torch.cuda.synchronize()
        ctx.decrypt_device(kids, d_ct.data_ptr(), 208, d_e.data_ptr(), 1024, mode=mode, stream=st.cuda_stream)
        st.synchronize()
        assert np.array_equal(d_e.cpu().numpy(), want)


def test_large_batch_round_trip(ctx, keys):
    """decrypt(encrypt(e)) == e on a few thousand items per set in one fused launch."""
    ct, kids, want = mixed_rows(keys, 2000, 63)
    got = ctx.decrypt(kids, ct, mode=cm.MODE_FUSED, ct_stride=208, e_stride=1024)
    assert np.array_equal(got, want)


def test_edge_inputs(ctx, keys):
    """All-zero and all-ones ciphertexts (zero / dense syndromes) match the oracle; an empty
    batch is a no-op."""
    for pid in PIDS:
        sk, kid = keys[pid]
        p = cm.params(pid)
        ct = np.zeros((3, p.ct_bytes), np.uint8)
        ct[1] = 0xFF
        ct[2, 0] = 1
        want = orc.decrypt(pid, sk, ct)
        for mode in (cm.MODE_PLAIN, cm.MODE_FUSED):
            assert np.array_equal(ctx.decrypt([kid] * 3, ct, mode=mode), want), (pid, mode)
    e = ctx.decrypt(np.zeros(0, np.int32), np.zeros((0, 208), np.uint8), ct_stride=208, e_stride=1024)
    assert e.shape == (0, 1024)


def test_argument_errors(ctx, keys):
    sk, kid = keys[348864]
    ct = np.zeros((2, 96), np.uint8)
    with pytest.raises(cm.CmdError) as ei:
        ctx.decrypt([kid, 99999], ct)
    assert ei.value.code == cm.CMD_ERR_KEY
    with pytest.raises(cm.CmdError) as ei:
        ctx.decrypt([kid, kid], ct, ct_stride=95, e_stride=436)
    assert ei.value.code == cm.CMD_ERR_ARG
    with pytest.raises(cm.CmdError) as ei:
        ctx.decrypt([kid, kid], ct, e_stride=435)
    assert ei.value.code == cm.CMD_ERR_ARG
    with pytest.raises(cm.CmdError) as ei:
        ctx.load_key(348864, sk[:-3])
    assert ei.value.code == cm.CMD_ERR_CRYPTO


def test_key_lifetime(ctx):
    """A live batch pins its keys; an unloaded key is rejected; slots are reused."""
    pid = 6960119
    sk = cm.gen_keypair(pid, 4242)
    kid = ctx.load_key(pid, sk)
    ct, e = cm.gen_encrypt(pid, sk, 5, 4)
    used0, budget = ctx.memory()
    b = cm.Batch(ctx, [kid] * 4, ct.shape[1], e.shape[1])
    with pytest.raises(cm.CmdError) as ei:
        ctx.unload_key(kid)
    assert ei.value.code == cm.CMD_ERR_KEY
    b.close()
    ctx.unload_key(kid)
    used1, _ = ctx.memory()
    assert used0 - used1 == cm.key_device_bytes(pid)
    with pytest.raises(cm.CmdError) as ei:
        ctx.decrypt([kid] * 4, ct)
    assert ei.value.code == cm.CMD_ERR_KEY
    with pytest.raises(cm.CmdError):
        ctx.unload_key(kid)
    kid2 = ctx.load_key(pid, sk)
    assert np.array_equal(ctx.decrypt([kid2] * 4, ct), e)
    ctx.unload_key(kid2)


def test_hbm_budget_admission():
    """Keys beyond the context's HBM budget are refused with CMD_ERR_BUDGET, and the
    context keeps working with the keys it has."""
    pid = 348864
    sks = cm.gen_keys(pid, 77, 4, threads=4)
    kb = cm.key_device_bytes(pid)
    c = cm.Context(0)
    used, _ = c.memory()
    c.close()
    with cm.Context(0, hbm_budget_bytes=used + 3 * kb + kb // 2) as c:
        loaded = []
        with pytest.raises(cm.CmdError) as ei:
            for _ in range(8):
                loaded.append(c.load_key(pid, sks[len(loaded) % 4]))
        assert ei.value.code == cm.CMD_ERR_BUDGET
        assert len(loaded) == 3
        c.unload_key(loaded.pop())
        u, b = c.memory()
        assert u <= b
        ct, e = cm.gen_encrypt(pid, sks[0], 3, 5)
        assert np.array_equal(c.decrypt([loaded[0]] * 5, ct), e)


def test_server_request_same_alone_and_in_mix(ctx, keys):
    """A request's error vector does not depend on what it was batched with: the same
    ciphertexts decode identically alone and inside a shuffled stream of all five sets."""
    pid = 8192128
    sk, kid = keys[pid]
    ct, e = cm.gen_encrypt(pid, sk, 91, 24)
    noise = []
    for q in PIDS:
        qsk, qkid = keys[q]
        qct, qe = cm.gen_encrypt(q, qsk, 92 + q % 7, 30)
        noise += [(qkid, c, x) for c, x in zip(qct, qe)]
    srv = cm.Server(ctx, max_batch=256, max_inflight=2, min_batch=1, max_wait_us=50.0)
    try:
        alone = []
        for c in ct:
            rid = srv.submit(kid, c)
            alone.append(srv.wait(rid, e.shape[1], 10e6))
        order = np.random.default_rng(4).permutation(len(noise) + len(ct))
        tagged = [("t", i) for i in range(len(ct))] + [("n", i) for i in range(len(noise))]
        ids = []
        for j in order:
            tag, i = tagged[j]
            if tag == "t":
                ids.append((tag, i, srv.submit(kid, ct[i])))
            else:
                ids.append((tag, i, srv.submit(noise[i][0], noise[i][1])))
        mixed = [None] * len(ct)
        for tag, i, rid in ids:
            if tag == "t":
                mixed[i] = srv.wait(rid, e.shape[1], 10e6)
            else:
                got = srv.wait(rid, noise[i][2].size, 10e6)
                assert np.array_equal(got, noise[i][2])
        st = srv.stats()
    finally:
        srv.close()
    assert np.array_equal(np.stack(alone), e)
    assert np.array_equal(np.stack(mixed), e)
    assert st["completed"] >= len(ids) and st["launches"] >= 1 and st["p99_us"] >= st["p50_us"] > 0


def test_server_poll_and_budget(ctx, keys):
    sk, kid = keys[460896]
    ct, e = cm.gen_encrypt(460896, sk, 7, 3)
    with pytest.raises(cm.CmdError) as ei:
        cm.Server(ctx, max_batch=4096, max_inflight=2, io_budget_bytes=1 << 20)
    assert ei.value.code == cm.CMD_ERR_BUDGET
    srv = cm.Server(ctx, max_batch=64, max_inflight=1)
    try:
        rids = [srv.submit(kid, c) for c in ct]
        with pytest.raises(cm.CmdError) as ei:
            srv.submit(123456, ct[0])
        assert ei.value.code == cm.CMD_ERR_KEY
        for rid, x in zip(rids, e):
            got = None
            for _ in range(100000):
                got = srv.poll(rid, x.size)
                if got is not None:
                    break
            assert got is not None and np.array_equal(got, x)
        with pytest.raises(cm.CmdError):
            srv.poll(rids[0], e.shape[1])      # a request is freed once its result is taken
    finally:
        srv.close()
    ctx_used = ctx.memory()[0]
    assert ctx_used > 0


def test_schedule_persistence_and_autotune(ctx, keys):
    pid = 6688128
    sk, kid = keys[pid]
    saved = {q: ctx.schedule(q) for q in PIDS}
    try:
        best, us = ctx.autotune(pid, [kid] * 2048, reps=1)
        assert us > 0 and best["tile_cts"] in (1, 2, 4, 8, 16, 32) and best["root_alg"] in (0, 1)
        assert ctx.schedule(pid) == best
        ct, e = cm.gen_encrypt(pid, sk, 17, 40)
        assert np.array_equal(ctx.decrypt([kid] * 40, ct), e)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "sched.json")
            ctx.save_schedules(path)
            doc = json.load(open(path))
            assert str(pid) in json.dumps(doc)
            ctx.set_schedule(pid, tile_cts=16, waves_per_group=1, root_alg=1 - best["root_alg"])
            assert ctx.schedule(pid) != best
            ctx.load_schedules(path)
            assert ctx.schedule(pid) == best
        with pytest.raises(cm.CmdError):
            ctx.set_schedule(pid, tile_cts=24)
    finally:
        for q, s in saved.items():
            ctx.set_schedule(q, **s)


def test_kernel_timing(ctx, keys):
    sk, kid = keys[348864]
    ct, e = cm.gen_encrypt(348864, sk, 3, 64)
    ctx.set_kernel_timing(True)
    try:
        ctx.decrypt([kid] * 64, ct)
        us, n = ctx.kernel_timing()
    finally:
        ctx.set_kernel_timing(False)
    assert n >= 1 and us > 0


def test_issue_probes_report_plausible_rates(ctx):
    """The VALU issue-rate and dependent-latency probes: rates are fractions of one instruction
    per clock, a second wavefront per SIMD raises the rate, and a readlane round trip through
    the scalar unit is slower than a DPP step, which is slower than a plain VALU step."""
    one, two = ctx.valu_issue_rate(1, 8), ctx.valu_issue_rate(2, 8)
    assert 0.02 < one < 1.5 and 0.02 < two < 1.5
    assert two > 1.3 * one
    xor, dpp, rdl = ctx.chain_latency()
    assert 0 < xor < dpp < rdl
    for bad in ((0, 8), (9, 8), (1, 3)):
        with pytest.raises(cm.CmdError):
            ctx.valu_issue_rate(*bad)


def _lbits_from_L(m, L, stride):
    out = np.zeros((m, stride), np.uint32)
    for k in range(m):
        bits = ((L >> k) & 1).astype(np.uint8)
        words = np.packbits(bits, bitorder="little")
        words = np.concatenate([words, np.zeros((-len(words)) % 4, np.uint8)]).view("<u4")
        out[k, :len(words)] = words
    return out


def _benes_from_cond(m, cond):
    count, words = 1 << m, 1 << (m - 5)
    out = np.zeros((2 * m - 1, words), np.uint32)
    c = np.arange(count)
    for k in range(2 * m - 1):
        s = k if k < m else 2 * m - 2 - k
        bits = np.unpackbits(cond[k * (count // 16):(k + 1) * (count // 16)], bitorder="little")
        lower = c[((c >> s) & 1) == 0]
        j = ((lower >> (s + 1)) << s) | (lower & ((1 << s) - 1))
        on = lower[bits[j] == 1]
        np.bitwise_or.at(out[k], on >> 5, (np.uint32(1) << (on & 31).astype(np.uint32)))
        if s >= 5:
            np.bitwise_or.at(out[k], (on + (1 << s)) >> 5, (np.uint32(1) << (on & 31).astype(np.uint32)))
    return out


@pytest.mark.parametrize("pid", PIDS)
def test_device_key_expansion_matches_host(ctx, pid):
    """cmd_key_load expands keys on the GPU: support, syndrome table, bit planes and Benes
    masks equal the host expansion (itself pinned against the reference build) for the shipped
    key and a generated one."""
    p = cm.params(pid)
    for sk in (orc.shipped_sk(pid), cm.gen_keypair(pid, 555)):
        kid = ctx.load_key(pid, sk)
        try:
            table, L, lbits, benes = ctx.export_key(kid)
        finally:
            ctx.unload_key(kid)
        g, L_host, T_host = cm.key_expand_host(pid, sk)
        assert np.array_equal(L, L_host)
        assert np.array_equal(table, T_host.reshape(p.sb, 2 * p.t))
        assert np.array_equal(lbits, _lbits_from_L(p.m, L_host, lbits.shape[1]))
        cond = np.asarray(sk[40 + 2 * p.t:40 + 2 * p.t + p.cond_bytes], np.uint8)
        assert np.array_equal(benes, _benes_from_cond(p.m, cond))


CLI = os.path.join(orc.ROOT, "cm_decrypt__amd", "cmdecrypt")


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_cli_drop_in_matches_reference_errorstream(tmp_path, mode):
    """The CLI invoked as decrypt_348864 in the reference's directory layout reads
    ct/sk_348864.bin and writes errorstream0_348864.bin byte-identical to the one the reference
    driver wrote for the same files (golden fixture)."""
    import subprocess
    gold = np.load(os.path.join(orc.GOLDEN, "golden_c1_348864.npz"))
    vec = tmp_path / "kem" / "test_vectors" / "Cipher_Sk"
    vec.mkdir(parents=True)
    gold["sk"].tofile(vec / "sk_348864.bin")
    gold["ct"].tofile(vec / "ct_348864.bin")
    work = tmp_path / "CPU" / "param" / "param348864"
    work.mkdir(parents=True)
    (tmp_path / "CPU" / "results" / "output").mkdir(parents=True)
    exe = work / "decrypt_348864"
    os.symlink(CLI, exe)
    r = subprocess.run([str(exe), str(len(gold["ct"])), "--mode", mode], cwd=work, capture_output=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    out = (tmp_path / "CPU" / "results" / "output" / "errorstream0_348864.bin").read_bytes()
    assert out == gold["stream"].tobytes()


def test_host_pipeline_chunks_pageable_and_pinned(ctx, keys):
    """Host-buffer calls larger than one chunk run through the three-slot pipeline (five
    chunks here, so slots are reused): results equal the encrypted errors for pageable and
    for cmd_host_alloc buffers, in fused and KEM modes, and the rows' padding is zero."""
    pid = 348864
    sk, kid = keys[pid]
    p = cm.params(pid)
    ct1, e1 = cm.gen_encrypt(pid, sk, 71, 4099)
    reps = 64  # 262336 items -> 5 chunks of about 52k
    ct = np.tile(ct1, (reps, 1))
    want = np.zeros((len(ct), 448), np.uint8)
    want[:, :p.e_bytes] = np.tile(e1, (reps, 1))
    print(f"host pipeline items={len(ct)}", flush=True)
    kids = np.full(len(ct), kid, np.int32)
    got = ctx.decrypt(kids, ct, e_stride=448)
    assert np.array_equal(got, want)
    h_ct = ctx.host_array(ct.shape)
    h_e = ctx.host_array((len(ct), 448))
    try:
        h_ct[:] = ct
        h_e[:] = 0xA5
        ctx.decrypt(kids, h_ct, e_stride=448, out=h_e)
        assert np.array_equal(h_e, want)
        # KEM through the pipeline: every key is SHAKE256(1 || e || c)
        k = ctx.kem_dec(kids[:70000], ct[:70000])
        import hashlib
        for i in (0, 1, 4098, 4099, 65535, 65536, 69999):
            assert k[i].tobytes() == hashlib.shake_256(b"\x01" + want[i, :p.e_bytes].tobytes()
                                                       + ct[i].tobytes()).digest(32)
    finally:
        ctx.host_free(h_ct)
        ctx.host_free(h_e)


def test_host_pipeline_mixed_sets_across_chunks(ctx, keys):
    """A shuffled five-set batch of three chunks through the host-buffer pipeline: each chunk
    has its own plan (sets, tiles, schedules) in its slot while the next is planned, and every
    row equals its encrypted error, pageable and pinned."""
    ct, kids, want = mixed_rows(keys, 2000, 4242)
    reps = 14  # 140000 items -> 3 chunks
    ct, kids, want = np.tile(ct, (reps, 1)), np.tile(kids, reps), np.tile(want, (reps, 1))
    order = np.random.default_rng(7).permutation(len(ct))
    ct, kids, want = ct[order], kids[order], want[order]
    print(f"mixed host pipeline items={len(ct)}", flush=True)
    assert np.array_equal(ctx.decrypt(kids, ct), want)
    h_ct = ctx.host_array(ct.shape)
    h_e = ctx.host_array(want.shape)
    try:
        h_ct[:] = ct
        h_e[:] = 0x5A
        ctx.decrypt(kids, h_ct, out=h_e)
        assert np.array_equal(h_e, want)
    finally:
        ctx.host_free(h_ct)
        ctx.host_free(h_e)


def _positions_rows(e, n, t):
    rows = np.full((len(e), t), 0xFFFF, np.uint16)
    for i, v in enumerate(e):
        p = np.flatnonzero(np.unpackbits(v, bitorder="little")[:n])[:t]
        rows[i, :len(p)] = p
    return rows


@pytest.mark.parametrize("pid", PIDS)
def test_positions_mode_vs_oracle(ctx, keys, pid):
    """CMD_MODE_POSITIONS rows equal the oracle's error vectors as positions: weight t,
    weight < t (the L_i = 0 quirk may add one) and random ciphertexts."""
    sk, kid = keys[pid]
    p = cm.params(pid)
    ct, _ = cm.gen_encrypt(pid, sk, 81, 40)
    low, _ = cm.gen_encrypt(pid, sk, 82, 5, weight=p.t - 7)
    rnd = np.random.default_rng(pid % 113).integers(0, 256, (8, p.ct_bytes), dtype=np.uint8)
    cts = np.vstack([ct, low, rnd])
    want = _positions_rows(orc.decrypt(pid, sk, cts), p.n, p.t)
    print(f"positions pid={pid} items={len(cts)}", flush=True)
    assert np.array_equal(ctx.positions([kid] * len(cts), cts), want)


def test_positions_mixed_batch_graph(ctx, keys):
    ct, kids, want_e = mixed_rows(keys, 19, 63)
    pid_of = {kid: pid for pid, (_, kid) in keys.items()}
    d_ct = torch.from_numpy(ct).cuda()
    d_p = torch.zeros((len(ct), 256), dtype=torch.uint8, device="cuda")
    b = cm.Batch(ctx, kids, 208, 256, mode=cm.MODE_POSITIONS, use_graph=True)
    try:
        for rep in range(3):
            d_p.zero_()
            b.run(d_ct.data_ptr(), d_p.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            got = d_p.cpu().numpy().view(np.uint16)
            for i in range(len(ct)):
                p = cm.params(pid_of[int(kids[i])])
                w = _positions_rows(want_e[i:i + 1, :p.e_bytes], p.n, p.t)[0]
                assert np.array_equal(got[i, :p.t], w), f"replay {rep} row {i}"
    finally:
        b.close()


@pytest.mark.parametrize("pid", PIDS)
def test_gpu_generator_matches_host_generator(ctx, pid):
    """cmd_gen_encrypt_device: per key the host generator's ciphertexts and errors, for
    seeded keys, the shipped key (its golden ciphertexts), and weight below t."""
    p = cm.params(pid)
    sks = [cm.gen_keypair(pid, 300 + k) for k in range(3)] + [orc.shipped_sk(pid)]
    kids = [ctx.load_key(pid, sk) for sk in sks]
    try:
        seeds = [11, 12, 13, 4242]
        print(f"gpu generator pid={pid}", flush=True)
        for weight, per in ((-1, 37), (p.t - 5, 6)):
            ct, e = ctx.gen_encrypt(kids, seeds, per, weight)
            for k, sk in enumerate(sks):
                want_ct, want_e = cm.gen_encrypt(pid, sk, seeds[k], per, weight=weight)
                assert np.array_equal(ct[k * per:(k + 1) * per], want_ct), f"key {k} weight {weight}"
                assert np.array_equal(e[k * per:(k + 1) * per], want_e)
        gold = np.load(os.path.join(orc.GOLDEN, f"golden_{pid}.npz"))
        ct, _ = ctx.gen_encrypt(kids[3:], [4242], 5)
        assert np.array_equal(ct, gold["ct"])
    finally:
        for k in kids:
            ctx.unload_key(k)


@pytest.mark.parametrize("pid", PIDS)
def test_gpu_keypairs_match_host_generator(ctx, pid):
    """cmd_gen_keys_device: the systematic-form test of every attempt on the GPU gives the host
    generator's secret keys byte for byte (about 3.5 attempts per key: random binary matrices
    are invertible with probability about 0.29, so redraws are exercised), and the keys load
    and decrypt."""
    n = 6
    want = cm.gen_keys(pid, 3300 + pid % 1000, n, threads=8)
    print(f"device keypairs pid={pid} keys={n}", flush=True)
    got = ctx.gen_keys(pid, 3300 + pid % 1000, n, threads=8)
    assert np.array_equal(got, want)
    kid = ctx.load_key(pid, got[-1])
    try:
        ct, e = cm.gen_encrypt(pid, got[-1], 5, 4)
        assert np.array_equal(ctx.decrypt([kid] * len(ct), ct)[:, :e.shape[1]], e)
    finally:
        ctx.unload_key(kid)
