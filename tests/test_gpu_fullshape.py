"""The benchmarked workloads at their full shapes, every row checked (the reference's contract
is an all-rows cmp of the errorstreams, GPU_Baseline/compare_errorstreams.sh:50):
  * mceliece8192128, 16 keys, 65536 ciphertexts round-robin, one fused launch replayed as a
    hipGraph: decrypt(encrypt(e)) == e on every row, and the oracle on a 256-row sample;
  * the five-set mix, 100000 ciphertexts over 64 keys per set, shuffled, one launch;
  * the serving loop over 1025 resident keys of the mix, 10^4 requests in bursts.
Keys come from the GPU key-pair generator (the host generator's bytes); ciphertexts from the
generator with exact-weight errors."""
import numpy as np
import pytest
import torch

import cm_decrypt__amd as cm
import _oracle as orc

pytestmark = pytest.mark.gpu
MIX = ((348864, 0.30), (460896, 0.20), (6688128, 0.20), (6960119, 0.15), (8192128, 0.15))


def test_headline_65536_x_16_keys_every_row(ctx):
    pid, n_keys, n = 8192128, 16, 65536
    sks = ctx.gen_keys(pid, 61, n_keys, threads=16)
    kids = [ctx.load_key(pid, s) for s in sks]
    try:
        key_of = np.arange(n) % n_keys
        ct = np.zeros((n, 208), np.uint8)
        e = np.zeros((n, 1024), np.uint8)
        cts, es = ctx.gen_encrypt(kids, [7000 + k for k in range(n_keys)], n // n_keys)
        per = n // n_keys
        for k in range(n_keys):
            rows = np.nonzero(key_of == k)[0]
            ct[rows] = cts[k * per:(k + 1) * per]
            e[rows] = es[k * per:(k + 1) * per]
        key_ids = np.array([kids[k] for k in key_of], np.int32)
        d_ct = torch.from_numpy(ct).cuda()
        d_e = torch.zeros((n, 1024), dtype=torch.uint8, device="cuda")
        print(f"headline shape: {n} ciphertexts, {n_keys} keys", flush=True)
        b = cm.Batch(ctx, key_ids, 208, 1024, mode=cm.MODE_FUSED, use_graph=True)
        try:
            for rep in range(2):
                d_e.zero_()
                b.run(d_ct.data_ptr(), d_e.data_ptr(), torch.cuda.current_stream().cuda_stream)
                torch.cuda.synchronize()
                got = d_e.cpu().numpy()
                bad = np.flatnonzero((got != e).any(axis=1))
                assert bad.size == 0, f"replay {rep}: {bad.size} rows differ, first {bad[:5]}"
        finally:
            b.close()
        sample = np.random.default_rng(3).choice(n, 256, replace=False)
        for k in range(n_keys):
            rows = sample[key_of[sample] == k]
            if rows.size:
                want = orc.decrypt(pid, sks[k], ct[rows])
                assert np.array_equal(got[rows], want), f"oracle sample, key {k}"
    finally:
        for k in kids:
            ctx.unload_key(k)


def test_mixed_100000_x_64_keys_per_set_every_row(ctx):
    n_items, per_set_keys = 100000, 64
    rng = np.random.default_rng(23)
    counts = {pid: int(round(n_items * w)) for pid, w in MIX}
    counts[MIX[0][0]] += n_items - sum(counts.values())
    ct = np.zeros((n_items, 208), np.uint8)
    e = np.zeros((n_items, 1024), np.uint8)
    kids_all = np.zeros(n_items, np.int32)
    order = rng.permutation(n_items)
    loaded = []
    pos = 0
    try:
        for pid, cnt in counts.items():
            sks = ctx.gen_keys(pid, 90 + pid % 1000, per_set_keys, threads=16)
            kids = [ctx.load_key(pid, s) for s in sks]
            loaded += kids
            per = -(-cnt // per_set_keys)
            cts, es = ctx.gen_encrypt(kids, [500 + k for k in range(per_set_keys)], per)
            p = cm.params(pid)
            rows = order[pos:pos + cnt]
            ct[rows, :p.ct_bytes] = cts[:cnt]
            e[rows, :p.e_bytes] = es[:cnt]
            kids_all[rows] = np.repeat(kids, per)[:cnt]
            pos += cnt
        print(f"mixed shape: {n_items} ciphertexts, {len(loaded)} keys", flush=True)
        d_ct = torch.from_numpy(ct).cuda()
        d_e = torch.zeros((n_items, 1024), dtype=torch.uint8, device="cuda")
        b = cm.Batch(ctx, kids_all, 208, 1024, mode=cm.MODE_FUSED, use_graph=True)
        try:
            b.run(d_ct.data_ptr(), d_e.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
        finally:
            b.close()
        got = d_e.cpu().numpy()
        bad = np.flatnonzero((got != e).any(axis=1))
        assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"
    finally:
        for k in loaded:
            ctx.unload_key(k)


def test_serving_1025_keys_10000_requests_every_result(ctx):
    n_keys, n_req = 1025, 10000
    per_set = {pid: max(1, int(round(n_keys * w))) for pid, w in MIX}
    templates_kid, templates_ct, templates_e = [], [], []
    loaded = []
    try:
        for pid, nk in per_set.items():
            sks = ctx.gen_keys(pid, 17 + pid % 1000, nk, threads=16)
            kids = [ctx.load_key(pid, s) for s in sks]
            loaded += kids
            cts, es = ctx.gen_encrypt(kids, [31 * k + 5 for k in range(nk)], 2)
            p = cm.params(pid)
            for i in range(len(cts)):
                r = np.zeros(208, np.uint8)
                r[:p.ct_bytes] = cts[i]
                w = np.zeros(1024, np.uint8)
                w[:p.e_bytes] = es[i]
                templates_kid.append(kids[i // 2])
                templates_ct.append(r)
                templates_e.append(w)
        t_kid = np.array(templates_kid, np.int32)
        t_ct = np.stack(templates_ct)
        t_e = np.stack(templates_e)
        picks = np.random.default_rng(5).choice(len(t_kid), n_req)
        print(f"serving shape: {len(loaded)} resident keys, {n_req} requests", flush=True)
        srv = cm.Server(ctx, max_batch=2048, max_inflight=4, min_batch=1, max_wait_us=100.0)
        try:
            rids = np.zeros(n_req, np.uint64)
            for i in range(0, n_req, 500):  # bursts that arrive while earlier ones run
                rids[i:i + 500] = srv.submit_batch(t_kid[picks[i:i + 500]], t_ct[picks[i:i + 500]], 208)
            got = srv.wait_batch(rids, 1024, 120e6)
            st = srv.stats()
        finally:
            srv.close()
        bad = np.flatnonzero((got != t_e[picks]).any(axis=1))
        assert bad.size == 0, f"{bad.size} results differ, first {bad[:5]}"
        assert st["failed"] == 0 and st["completed"] == n_req
    finally:
        for k in loaded:
            ctx.unload_key(k)
