"""GPU parity: every kernel path against the oracle (C restatement of the reference
decoder), bit for bit, on seeded inputs.  Runs through the C ABI of libcmdecrypt.so."""
import numpy as np
import pytest

import cm_decrypt__amd as cm
import _oracle as orc

pytestmark = pytest.mark.gpu
PIDS = list(cm.PARAM_IDS)


@pytest.fixture(scope="module")
def keys(ctx):
    out = {}
    for pid in PIDS:
        sk = cm.gen_keypair(pid, 7000 + pid % 1000)
        out[pid] = (sk, ctx.load_key(pid, sk))
    return out


@pytest.fixture(scope="module")
def shipped(ctx):
    out = {}
    for pid in PIDS:
        sk = orc.shipped_sk(pid)
        out[pid] = (sk, ctx.load_key(pid, sk))
    return out


def rand_ct(pid, count, seed):
    p = cm.params(pid)
    return np.random.default_rng(seed).integers(0, 256, (count, p.ct_bytes), dtype=np.uint8)


def inputs(pid, sk, n_valid, n_rand, seed):
    ct, e = cm.gen_encrypt(pid, sk, seed, n_valid)
    ct = np.concatenate([ct, rand_ct(pid, n_rand, seed + 1)])
    return ct, e


@pytest.mark.parametrize("pid", PIDS)
def test_stage_synd(ctx, keys, pid):
    sk, kid = keys[pid]
    ct, _ = inputs(pid, sk, 6, 6, 11)
    g, L, T = orc.key_expand(pid, sk)
    got = ctx.synd(kid, ct)
    want = np.stack([orc.synd(pid, T, c) for c in ct])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("pid", PIDS)
def test_stage_bm(ctx, keys, pid):
    sk, kid = keys[pid]
    ct, _ = inputs(pid, sk, 6, 6, 12)
    g, L, T = orc.key_expand(pid, sk)
    s = np.stack([orc.synd(pid, T, c) for c in ct])
    s_rand = np.random.default_rng(5).integers(0, 1 << cm.params(pid).m, (4, s.shape[1]), dtype=np.uint16)
    s = np.concatenate([s, s_rand, np.zeros((1, s.shape[1]), np.uint16)])
    got = ctx.bm(pid, s)
    want = np.stack([orc.bm(pid, x) for x in s])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("pid", PIDS)
def test_stage_root(ctx, keys, pid):
    sk, kid = keys[pid]
    ct, _ = inputs(pid, sk, 4, 4, 13)
    g, L, T = orc.key_expand(pid, sk)
    loc = np.stack([orc.bm(pid, orc.synd(pid, T, c)) for c in ct])
    got = ctx.root(kid, loc)
    want = np.stack([orc.error_vector(pid, orc.root(pid, f, L)) for f in loc])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", [cm.MODE_PLAIN, cm.MODE_FUSED])
@pytest.mark.parametrize("pid", PIDS)
def test_decrypt_one_key(ctx, keys, pid, mode):
    sk, kid = keys[pid]
    ct, e = inputs(pid, sk, 37, 8, 21)          # 45 items: one full + one partial tile
    got = ctx.decrypt([kid] * len(ct), ct, mode=mode)
    want = orc.decrypt(pid, sk, ct)
    assert np.array_equal(got[:, :want.shape[1]], want)
    assert np.array_equal(got[:37, :e.shape[1]], e)   # valid ciphertexts recover the error exactly


@pytest.mark.parametrize("mode", [cm.MODE_PLAIN, cm.MODE_FUSED])
def test_decrypt_shipped_keys(ctx, shipped, mode):
    for pid in PIDS:
        sk, kid = shipped[pid]
        ct, e = cm.gen_encrypt(pid, sk, 31, 5)
        got = ctx.decrypt([kid] * 5, ct, mode=mode)
        assert np.array_equal(got, e), pid


@pytest.mark.parametrize("mode", [cm.MODE_PLAIN, cm.MODE_FUSED])
def test_decrypt_mixed_sets_and_keys(ctx, keys, shipped, mode):
    rng = np.random.default_rng(3)
    rows, key_ids, want = [], [], []
    for pid in PIDS:
        for (sk, kid), n in ((keys[pid], 19), (shipped[pid], 7)):
            ct, e = inputs(pid, sk, n, 2, 40 + n)
            w = orc.decrypt(pid, sk, ct)
            for c, x in zip(ct, w):
                r = np.zeros(208, np.uint8)
                r[:c.size] = c
                rows.append(r)
                key_ids.append(kid)
                wx = np.zeros(1024, np.uint8)
                wx[:x.size] = x
                want.append(wx)
    order = rng.permutation(len(rows))
    ct = np.stack(rows)[order]
    kids = np.array(key_ids, np.int32)[order]
    want = np.stack(want)[order]
    got = ctx.decrypt(kids, ct, mode=mode, ct_stride=208, e_stride=1024)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("pid", [348864, 6960119])
def test_key_runs_inside_tiles(ctx, pid):
    """Tiles holding ciphertexts of several keys (items are planned in key order): runs of one
    key share a table pass, tiles of many short runs read ciphertext by ciphertext, a tile of at
    most three ciphertexts takes the scalar path; every row against the oracle, on distinct keys
    so that a table taken from the wrong key shows."""
    sks = cm.gen_keys(pid, 91, 9, threads=8)
    kids = [ctx.load_key(pid, s) for s in sks]
    # 32-wide tiles: [1,2,3,5,8,13] (six runs), [1,30,1] (three), [3] (scalar); then one tile
    # of nine single-ciphertext runs; then [16,16] (two runs of half a tile)
    layouts = [[1, 2, 3, 5, 8, 13, 1, 30, 4], [1] * 9, [16, 16]]
    saved = ctx.schedule(pid)
    try:
        ctx.set_tile_fit(False)
        ctx.set_schedule(pid, tile_cts=32, waves_per_group=1, groups_per_cu=8, root_alg=1)
        for li, counts in enumerate(layouts):
            ct_rows, kk, want = [], [], []
            for k, n in enumerate(counts):
                ct, _ = cm.gen_encrypt(pid, sks[k], 300 + 10 * li + k, n)
                want.append(orc.decrypt(pid, sks[k], ct))
                ct_rows.append(ct)
                kk += [kids[k]] * len(ct)
            ct = np.concatenate(ct_rows)
            want = np.concatenate(want)
            order = np.random.default_rng(li).permutation(len(ct))
            got = ctx.decrypt(np.array(kk, np.int32)[order], ct[order], mode=cm.MODE_FUSED)
            assert np.array_equal(got, want[order]), f"layout {counts}"
    finally:
        ctx.set_tile_fit(True)
        ctx.set_schedule(pid, **saved)
        for k in kids:
            ctx.unload_key(k)


@pytest.mark.parametrize("pid", [348864, 460896])
def test_random_key_runs_fitted_tiles(ctx, pid):
    """Random run lengths over twelve keys, valid and random ciphertexts, shuffled, with the
    planner's own tile widths (fitted to the batch): every row against the oracle."""
    rng = np.random.default_rng(pid % 1000)
    sks = cm.gen_keys(pid, 57, 12, threads=8)
    kids = [ctx.load_key(pid, s) for s in sks]
    try:
        rows, kk, want = [], [], []
        for k in range(12):
            n = int(rng.integers(1, 60))
            ct, _ = inputs(pid, sks[k], n, int(rng.integers(0, 3)), 700 + k)
            want.append(orc.decrypt(pid, sks[k], ct))
            rows.append(ct)
            kk += [kids[k]] * len(ct)
        ct = np.concatenate(rows)
        want = np.concatenate(want)
        order = rng.permutation(len(ct))
        got = ctx.decrypt(np.array(kk, np.int32)[order], ct[order], mode=cm.MODE_FUSED)
        assert np.array_equal(got, want[order])
    finally:
        for k in kids:
            ctx.unload_key(k)


@pytest.mark.parametrize("pid", PIDS)
def test_weight_below_t(ctx, keys, pid):
    """w < t leaves a factor x^(t-w) in the reversed locator: a support element equal to 0
    is reported as an error too (the reference's behaviour, reproduced bit for bit)."""
    sk, kid = keys[pid]
    t = cm.params(pid).t
    ct, e = cm.gen_encrypt(pid, sk, 77, 6, weight=t - 5)
    want = orc.decrypt(pid, sk, ct)
    for mode in (cm.MODE_PLAIN, cm.MODE_FUSED):
        got = ctx.decrypt([kid] * len(ct), ct, mode=mode)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("wpg,gpc", [(1, 1), (2, 1), (1, 3)])
def test_more_tiles_than_wavefronts(ctx, keys, wpg, gpc):
    """Launches with more tiles than the grid has wavefronts: each wavefront's first tile comes
    from the XCD-major static assignment, the rest from the counter past the grid; every tile
    is decrypted once (grids of n_CU x gpc workgroups of one or two wavefronts, so 256 to 768
    wavefronts for 625 tiles)."""
    pid = 348864
    sk, kid = keys[pid]
    ct, e = cm.gen_encrypt(pid, sk, 4321, 20000)  # 625 tiles of 32
    saved = ctx.schedule(pid)
    try:
        ctx.set_tile_fit(False)
        ctx.set_schedule(pid, tile_cts=32, waves_per_group=wpg, waves_per_ct=1, groups_per_cu=gpc, root_alg=1)
        got = ctx.decrypt([kid] * len(ct), ct, mode=cm.MODE_FUSED)
    finally:
        ctx.set_tile_fit(True)
        ctx.set_schedule(pid, **saved)
    assert np.array_equal(got, e)


@pytest.mark.parametrize("alg", [0, 1])
@pytest.mark.parametrize("tile,wpg,gpc", [(16, 1, 2), (32, 2, 2), (32, 4, 1), (32, 1, 8), (16, 2, 4), (32, 4, 2),
                                          (4, 2, 4), (1, 4, 1)])
def test_fused_schedules_agree(ctx, keys, alg, tile, wpg, gpc):
    """Every fused-kernel variant (Horner or FFT root search, tile width, wavefronts per group,
    one- or two-wavefront-per-SIMD build) gives the oracle's error vectors on a mixed batch."""
    rows, kids, want = [], [], []
    for pid in PIDS:
        sk, kid = keys[pid]
        ct, _ = inputs(pid, sk, 9, 3, 500 + pid % 97)
        w = orc.decrypt(pid, sk, ct)
        for c, x in zip(ct, w):
            r = np.zeros(208, np.uint8)
            r[:c.size] = c
            rows.append(r)
            kids.append(kid)
            wx = np.zeros(1024, np.uint8)
            wx[:x.size] = x
            want.append(wx)
    saved = {pid: ctx.schedule(pid) for pid in PIDS}
    try:
        ctx.set_tile_fit(False)  # the schedule's own tile width
        for pid in PIDS:
            ctx.set_schedule(pid, tile_cts=tile, waves_per_group=wpg, groups_per_cu=gpc, root_alg=alg)
        got = ctx.decrypt(np.array(kids, np.int32), np.stack(rows), mode=cm.MODE_FUSED, ct_stride=208, e_stride=1024)
        ctx.set_tile_fit(True)
        fitted = ctx.decrypt(np.array(kids, np.int32), np.stack(rows), mode=cm.MODE_FUSED, ct_stride=208,
                             e_stride=1024)
    finally:
        ctx.set_tile_fit(True)
        for pid, s in saved.items():
            ctx.set_schedule(pid, **s)
    want = np.stack(want)
    bad = [i for i in range(len(rows)) if not np.array_equal(got[i], want[i])]
    detail = [(i, int(np.nonzero(got[i] != want[i])[0][0])) for i in bad[:6]]
    assert not bad, f"{len(bad)} rows differ (row, first byte): {detail}"
    assert np.array_equal(fitted, want), "narrowed tiles differ"


@pytest.mark.parametrize("tile,wpg,gpc", [(1, 4, 1), (1, 2, 2), (2, 4, 2)])
def test_multiwave_workgroups_full_grid(ctx, keys, tile, wpg, gpc):
    """Launches with at least as many tiles as resident wavefronts keep the schedule's
    wavefronts per workgroup (smaller ones run one per workgroup): every wavefront of a
    multi-wavefront group, each on its own slice of LDS, recovers the generated errors."""
    pid = 348864
    sk, kid = keys[pid]
    n = tile * 256 * wpg * gpc + 77  # more tiles than the grid has wavefronts (256 CUs)
    ct, e = cm.gen_encrypt(pid, sk, 901 + tile, n)
    saved = ctx.schedule(pid)
    try:
        ctx.set_tile_fit(False)
        ctx.set_schedule(pid, tile_cts=tile, waves_per_group=wpg, groups_per_cu=gpc, root_alg=1)
        got = ctx.decrypt([kid] * n, ct, mode=cm.MODE_FUSED)
    finally:
        ctx.set_tile_fit(True)
        ctx.set_schedule(pid, **saved)
    assert np.array_equal(got[:, :e.shape[1]], e)


@pytest.mark.parametrize("pid,K", [(348864, 63), (8192128, 63), (8192128, 127), (6688128, 127)])
def test_late_first_syndrome(ctx, keys, pid, K):
    """Syndromes zero below index K put the first nonzero discrepancy on a 64-coefficient
    block boundary (B = x^(K+1)); every path, including the bit-sliced and the small-tile
    scalar Berlekamp-Massey, matches the oracle."""
    from _gf2 import late_syndrome_cts
    sk, kid = keys[pid]
    ct = late_syndrome_cts(pid, sk, K, 5, seed=K)
    want = orc.decrypt(pid, sk, ct)
    _, _, T = orc.key_expand(pid, sk)
    s = np.stack([orc.synd(pid, T, c) for c in ct])
    assert np.array_equal(ctx.bm(pid, s), np.stack([orc.bm(pid, x) for x in s]))
    saved = ctx.schedule(pid)
    try:
        ctx.set_tile_fit(False)
        for tile in (32, 1):
            ctx.set_schedule(pid, tile_cts=tile)
            for mode in (cm.MODE_PLAIN, cm.MODE_FUSED):
                got = ctx.decrypt([kid] * len(ct), ct, mode=mode)
                assert np.array_equal(got, want), (tile, mode)
    finally:
        ctx.set_tile_fit(True)
        ctx.set_schedule(pid, **saved)
