"""cmd_decrypt_device_noalloc through the C ABI: the allocation-free fused decryption gives the
same error rows as cmd_decrypt_device (and the generated errors) on two calls in a row, the
second smaller than the first, and charges no device memory to the context."""
import ctypes

import numpy as np
import pytest
import torch

import cm_decrypt__amd as cm

pytestmark = pytest.mark.gpu
SETS = (8192128, 6960119, 348864)


@pytest.fixture(scope="module")
def noalloc_keys(ctx):
    out = {}
    for pid in SETS:
        sk = cm.gen_keypair(pid, 4400 + pid % 1000)
        out[pid] = (sk, ctx.load_key(pid, sk))
    yield out
    for _, kid in out.values():
        ctx.unload_key(kid)


def rows(keys, per_set, seed):
    ct, kid, want = [], [], []
    for pid in SETS:
        sk, k = keys[pid]
        c, e = cm.gen_encrypt(pid, sk, seed + pid % 97, per_set)
        for a, b in zip(c, e):
            r = np.zeros(208, np.uint8)
            r[:a.size] = a
            w = np.zeros(1024, np.uint8)
            w[:b.size] = b
            ct.append(r)
            kid.append(k)
            want.append(w)
    order = np.random.default_rng(seed).permutation(len(ct))
    return np.stack(ct)[order], np.array(kid, np.int32)[order], np.stack(want)[order]


def noalloc(ctx, kids, d_ct, d_e):
    kids = np.ascontiguousarray(kids, np.int32)
    return cm.lib.cmd_decrypt_device_noalloc(ctx.handle, kids.size, kids.ctypes.data_as(ctypes.c_void_p),
                                             d_ct.data_ptr(), 208, d_e.data_ptr(), 1024, None)


def test_noalloc_matches_decrypt_device(ctx, noalloc_keys):
    used0, _ = ctx.memory()
    for call, per_set in enumerate((1700, 230)):
        ct, kids, want = rows(noalloc_keys, per_set, 71 + call)
        d_ct = torch.from_numpy(ct).cuda()
        d_ref = torch.zeros((len(ct), 1024), dtype=torch.uint8, device="cuda")
        d_na = torch.zeros((len(ct), 1024), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        print(f"noalloc call {call}: {len(ct)} items of {SETS}", flush=True)
        ctx.decrypt_device(kids, d_ct.data_ptr(), 208, d_ref.data_ptr(), 1024)
        before, _ = ctx.memory()
        rc = noalloc(ctx, kids, d_ct, d_na)
        assert rc == cm.CMD_OK, cm.lib.cmd_last_error().decode()
        after, _ = ctx.memory()
        assert after == before  # nothing charged to the context's device budget
        ref, got = d_ref.cpu().numpy(), d_na.cpu().numpy()
        assert np.array_equal(ref, want)
        assert np.array_equal(got, ref), f"call {call}: allocation-free rows differ from cmd_decrypt_device"
    assert ctx.memory()[0] == used0


def test_noalloc_rejects_oversized_call(ctx, noalloc_keys):
    _, kid = noalloc_keys[348864]
    kids = np.full(cm.NOALLOC_MAX_ITEMS + 1, kid, np.int32)
    # checked on the host before any device work: the buffers are never touched
    rc = cm.lib.cmd_decrypt_device_noalloc(ctx.handle, kids.size, kids.ctypes.data_as(ctypes.c_void_p),
                                           ctypes.c_void_p(16), 96, ctypes.c_void_p(16), 512, None)
    assert rc == cm.CMD_ERR_ARG
    assert "at most" in cm.lib.cmd_last_error().decode()
