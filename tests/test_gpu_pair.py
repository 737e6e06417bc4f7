"""Two wavefronts per tile (schedule waves_per_ct = 2): the paired fused kernel splits a tile's
syndrome by table column and its Berlekamp-Massey recurrence by lane block (the partial
discrepancies meet in LDS every iteration).  Every path against the oracle, bit for bit: all
five sets mixed in one launch (valid and random ciphertexts), tile widths down to the
one-ciphertext scalar path, first nonzero discrepancies on the 64-coefficient block
boundaries, tiles of many key runs, weight below t, and queues deeper than the grid."""
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
        sk = cm.gen_keypair(pid, 8100 + pid % 1000)
        out[pid] = (sk, ctx.load_key(pid, sk))
    yield out
    for _, kid in out.values():
        ctx.unload_key(kid)


class paired:
    """Pair schedule on the given sets for the duration of a block (tile widths as set)."""

    def __init__(self, ctx, pids, tile=32, gpc=8, fit=False):
        self.ctx, self.pids, self.tile, self.gpc, self.fit = ctx, pids, tile, gpc, fit

    def __enter__(self):
        self.saved = {pid: self.ctx.schedule(pid) for pid in self.pids}
        self.ctx.set_tile_fit(self.fit)
        for pid in self.pids:
            self.ctx.set_schedule(pid, tile_cts=self.tile, waves_per_group=2, waves_per_ct=2,
                                  groups_per_cu=self.gpc, root_alg=1)

    def __exit__(self, *exc):
        self.ctx.set_tile_fit(True)
        for pid, s in self.saved.items():
            self.ctx.set_schedule(pid, **s)


def mixed(keys, n_valid, n_rand, seed):
    rows, kids, want = [], [], []
    for pid in PIDS:
        sk, kid = keys[pid]
        p = cm.params(pid)
        ct, _ = cm.gen_encrypt(pid, sk, seed + pid % 89, n_valid)
        rnd = np.random.default_rng(seed + pid % 7).integers(0, 256, (n_rand, p.ct_bytes), dtype=np.uint8)
        ct = np.concatenate([ct, rnd])
        w = orc.decrypt(pid, sk, ct)
        for c, x in zip(ct, w):
            r = np.zeros(208, np.uint8)
            r[:c.size] = c
            wx = np.zeros(1024, np.uint8)
            wx[:x.size] = x
            rows.append(r)
            kids.append(kid)
            want.append(wx)
    order = np.random.default_rng(seed).permutation(len(rows))
    return np.stack(rows)[order], np.array(kids, np.int32)[order], np.stack(want)[order]


@pytest.mark.parametrize("tile", [32, 16, 4, 2, 1])
def test_pair_mixed_sets_vs_oracle(ctx, keys, tile):
    """All five sets in one paired launch; tiles of <= 3 ciphertexts take wave 0's scalar BM."""
    ct, kids, want = mixed(keys, 40, 9, 900 + tile)
    print(f"pair tile={tile} items={len(ct)}", flush=True)
    with paired(ctx, PIDS, tile=tile):
        got = ctx.decrypt(kids, ct, mode=cm.MODE_FUSED, ct_stride=208, e_stride=1024)
    bad = [i for i in range(len(ct)) if not np.array_equal(got[i], want[i])]
    assert not bad, f"{len(bad)} rows differ, first {bad[:5]}"


@pytest.mark.parametrize("pid,K", [(348864, 63), (8192128, 63), (8192128, 127), (6688128, 127),
                                   (460896, 63), (6960119, 63)])
def test_pair_late_first_syndrome(ctx, keys, pid, K):
    """First nonzero discrepancy at iteration K (B = x^(K+1)): the carry between the two
    waves' lane blocks and the top coefficient's entry are exercised at their boundary."""
    from _gf2 import late_syndrome_cts
    sk, kid = keys[pid]
    ct = late_syndrome_cts(pid, sk, K, 5, seed=K + 1)
    want = orc.decrypt(pid, sk, ct)
    print(f"pair late syndrome pid={pid} K={K}", flush=True)
    for tile in (32, 1):
        with paired(ctx, [pid], tile=tile):
            got = ctx.decrypt([kid] * len(ct), ct, mode=cm.MODE_FUSED)
        assert np.array_equal(got, want), f"tile {tile}"


@pytest.mark.parametrize("pid", [8192128, 6960119])
def test_pair_key_runs_inside_tiles(ctx, pid):
    """Tiles spanning several keys: runs share a column pass, tiles of many short runs are
    scanned ciphertext by ciphertext (distinct keys, so a wrong table shows)."""
    sks = cm.gen_keys(pid, 93, 9, threads=8)
    kids = [ctx.load_key(pid, s) for s in sks]
    try:
        for li, counts in enumerate([[1, 2, 3, 5, 8, 13, 1, 30, 4], [1] * 9, [16, 16]]):
            ct_rows, kk, want = [], [], []
            for k, n in enumerate(counts):
                ct, _ = cm.gen_encrypt(pid, sks[k], 400 + 10 * li + k, n)
                want.append(orc.decrypt(pid, sks[k], ct))
                ct_rows.append(ct)
                kk += [kids[k]] * len(ct)
            ct = np.concatenate(ct_rows)
            want = np.concatenate(want)
            order = np.random.default_rng(li).permutation(len(ct))
            print(f"pair key runs pid={pid} layout={counts}", flush=True)
            with paired(ctx, [pid]):
                got = ctx.decrypt(np.array(kk, np.int32)[order], ct[order], mode=cm.MODE_FUSED)
            assert np.array_equal(got, want[order]), f"layout {counts}"
    finally:
        for k in kids:
            ctx.unload_key(k)


@pytest.mark.parametrize("pid", PIDS)
def test_pair_weight_below_t(ctx, keys, pid):
    sk, kid = keys[pid]
    t = cm.params(pid).t
    ct, _ = cm.gen_encrypt(pid, sk, 78, 7, weight=t - 5)
    want = orc.decrypt(pid, sk, ct)
    with paired(ctx, [pid], tile=8):
        got = ctx.decrypt([kid] * len(ct), ct, mode=cm.MODE_FUSED)
    assert np.array_equal(got, want)


def test_pair_queue_deeper_than_grid(ctx, keys):
    """More tiles than resident pairs (each pair takes several tiles from the queue):
    decrypt(encrypt(e)) == e on every row."""
    pid = 8192128
    sk, kid = keys[pid]
    n = 256 * 2 * 4 * 3 + 41  # tile 4, two pairs per CU: three rounds of the grid and a tail
    ct, e = cm.gen_encrypt(pid, sk, 905, n)
    print(f"pair deep queue items={n}", flush=True)
    with paired(ctx, [pid], tile=4, gpc=2):
        got = ctx.decrypt([kid] * n, ct, mode=cm.MODE_FUSED)
    assert np.array_equal(got, e)
