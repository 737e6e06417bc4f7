// Host-buffer decryption pipeline and pinned host memory.
//
// cmd_decrypt on host buffers (fused, KEM and positions modes) cuts the items into chunks that flow
// through kIoSlots slots.  Each slot owns device buffers for one chunk and pinned staging.  All
// copies go on one copy stream and all kernels on one kernel stream, ordered by events: the copy
// in of one chunk and the copy out of another overlap the kernels of a third (the reference's
// driver runs them back to back per BATCH_SIZE chunk,
// GPU_Baseline/param/param8192128/src/cuda/Decrypt.cu:177-324).  Chunk i enqueues its copy
// in, then chunk i-1's copy out (which waits for chunk i-1's kernels), then its own kernels:
// the copy stream holds in(0), in(1), out(0), in(2), out(1), ..., and each copy in lands before
// the kernel stream needs it, one chunk's kernels later.  The two streams may share one
// hardware queue (the runtime caps the queues per process), and there every event record and
// cross-stream wait becomes a packet that holds back all packets behind it until it resolves:
// so the event that marks chunk i-1's copy out done is recorded only after chunk i's kernels
// are enqueued, and no wait is placed ahead of a kernel that it would delay.  The kernels of
// different chunks never run at once, so no persistent launch shares the GPU with another
// chunk's launch or fill.  Buffers the caller got from cmd_host_alloc are
// already pinned: they are copied to and from the device directly; other (pageable) buffers go
// through the slot's staging with multi-threaded copies.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.h"
#include "trace.h"

using namespace cmd;

namespace {

constexpr size_t kChunkItems = 65536;   // items per chunk: one full-GPU fused launch
constexpr size_t kMinChunkItems = 32768;

// memcpy over up to 8 threads for large copies.
void copy_par(void *dst, const void *src, size_t bytes)
{
    const size_t kPer = (size_t)4 << 20;
    const size_t nt = std::min<size_t>(8, std::max<size_t>(1, bytes / kPer));
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + nt - 1) / nt;
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) {
        const size_t a = t * per, b = std::min(bytes, a + per);
        if (a < b) pool.emplace_back([=] { std::memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    std::memcpy(dst, src, std::min(bytes, per));
    for (auto &th : pool) th.join();
}

bool is_pinned(const cmd_ctx *ctx, const void *p, size_t bytes)
{
    const uintptr_t a = (uintptr_t)p;
    for (const auto &r : ctx->host_pinned)
        if (a >= r.first && a + bytes <= r.first + r.second) return true;
    return false;
}

int host_reserve(void *&ptr, size_t &cap, size_t bytes)
{
    if (cap >= bytes && ptr) return CMD_OK;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    hipError_t e = hipHostMalloc(&ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, "hipHostMalloc");
    }
    cap = bytes;
    return CMD_OK;
}

// Enqueues a slot's copy out on the copy stream, behind its kernels; io_mark_done records its
// completion event.  The two are apart so that the record can go behind the next chunk's
// kernels (see above).
int io_copy_out(cmd_ctx *ctx, IoSlot &s)
{
    if (!s.out_pending) return CMD_OK;
    s.out_pending = false;
    hipError_t he = hipStreamWaitEvent(ctx->io_copy, s.k_done, 0);
    if (he == hipSuccess) he = hipMemcpyAsync(s.dst, s.d_out.ptr, s.out_bytes, hipMemcpyDeviceToHost, ctx->io_copy);
    if (he != hipSuccess) {
        (void)hipStreamSynchronize(ctx->io_kern);
        (void)hipStreamSynchronize(ctx->io_copy);
        s.busy = false;
        return hip_fail(he, "pipeline copy out");
    }
    s.mark_pending = true;
    return CMD_OK;
}

int io_mark_done(cmd_ctx *ctx, IoSlot &s)
{
    if (!s.mark_pending) return CMD_OK;
    s.mark_pending = false;
    const hipError_t he = hipEventRecord(s.done, ctx->io_copy);
    if (he != hipSuccess) {
        (void)hipStreamSynchronize(ctx->io_copy);
        s.busy = false;
        return hip_fail(he, "pipeline copy out");
    }
    return CMD_OK;
}

// Waits for a slot's chunk (enqueuing its copy out first if still pending) and copies staged
// results to the caller.
int io_finish(cmd_ctx *ctx, IoSlot &s, uint8_t *out, size_t out_stride)
{
    if (!s.busy) return CMD_OK;
    int rc = io_copy_out(ctx, s);
    if (!rc) rc = io_mark_done(ctx, s);
    if (rc) return rc;
    s.busy = false;
    CMD_HIP(hipEventSynchronize(s.done));
    if (s.timed) {
        float ms = 0.f;
        CMD_HIP(hipEventElapsedTime(&ms, s.t0, s.t1));
        ctx->timing_us += ms * 1000.0;
        ctx->timing_n++;
        s.timed = false;
    }
    if (s.staged_out) copy_par(out + s.first * out_stride, s.h_out, s.count * out_stride);
    return CMD_OK;
}

}  // namespace

namespace cmd {

void io_release(cmd_ctx *ctx)
{
    for (hipStream_t st : {ctx->io_copy, ctx->io_kern})
        if (st) (void)hipStreamSynchronize(st);
    for (IoSlot &s : ctx->io) {
        for (DevBuf *b : {&s.d_ct, &s.d_out, &s.d_plan, &s.d_kem_e, &s.d_ok, &s.d_msg}) dev_free(ctx, *b);
        if (s.h_ct) (void)hipHostFree(s.h_ct);
        if (s.h_out) (void)hipHostFree(s.h_out);
        if (s.h_plan) (void)hipHostFree(s.h_plan);
        for (hipEvent_t ev : {s.in_done, s.k_done, s.done, s.t0, s.t1})
            if (ev) (void)hipEventDestroy(ev);
    }
    ctx->io.clear();
    for (hipStream_t *st : {&ctx->io_copy, &ctx->io_kern})
        if (*st) (void)hipStreamDestroy(*st);
    ctx->io_copy = ctx->io_kern = nullptr;
    for (const auto &r : ctx->host_pinned) (void)hipHostFree((void *)r.first);
    ctx->host_pinned.clear();
}

int io_decrypt(cmd_ctx *ctx, int mode, size_t count, const int32_t *key_ids, const uint8_t *ct, size_t ct_stride,
               uint8_t *out, size_t out_stride)
{
    // validate every item first: a bad key must not leave half the output written
    TraceRange tv("pipeline validate");
    // one pass over the items through a per-key table of set bits (0: not loaded)
    std::vector<uint8_t> set_bit(ctx->keys.size(), 0);
    for (size_t k = 0; k < set_bit.size(); k++)
        if (ctx->keys[k].used && ctx->keys[k].pidx >= 0) set_bit[k] = (uint8_t)(1u << ctx->keys[k].pidx);
    unsigned sets = 0;
    for (size_t i = 0; i < count; i++) {
        const uint32_t k = (uint32_t)key_ids[i];
        const unsigned b = k < set_bit.size() ? set_bit[k] : 0u;
        if (!b) return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)key_ids[i]);
        sets |= b;
    }
    size_t max_ct = 0, max_e = 0, max_t = 0;
    for (int pidx = 0; pidx < kNumParams; pidx++)
        if (sets & (1u << pidx)) {
            max_ct = std::max(max_ct, (size_t)kParams[pidx].synd_bytes);
            max_e = std::max(max_e, (size_t)kParams[pidx].n / 8);
            max_t = std::max(max_t, (size_t)kParams[pidx].t);
        }
    const size_t need_out = mode == CMD_MODE_KEM ? CMD_KEM_KEY_BYTES : mode == CMD_MODE_POSITIONS ? 2 * max_t : max_e;
    if (mode == CMD_MODE_POSITIONS && (out_stride & 1u))
        return set_error(CMD_ERR_ARG, "position rows need an even stride (%zu)", out_stride);
    if (ct_stride < max_ct || out_stride < need_out)
        return set_error(CMD_ERR_ARG, "strides (%zu, %zu) below item sizes (%zu, %zu)", ct_stride, out_stride, max_ct,
                         need_out);
    tv.end();
    if (ctx->io.empty()) {
        // all streams and slots or none: a failure leaves no half-made pipeline behind
        std::vector<IoSlot> slots(kIoSlots);
        hipStream_t st[2] = {nullptr, nullptr};
        hipError_t he = hipSuccess;
        for (hipStream_t &x : st)
            if (he == hipSuccess) he = hipStreamCreateWithFlags(&x, hipStreamNonBlocking);
        for (IoSlot &s : slots) {
            for (hipEvent_t *ev : {&s.in_done, &s.k_done, &s.done})
                if (he == hipSuccess) he = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (he == hipSuccess) he = hipEventCreate(&s.t0);
            if (he == hipSuccess) he = hipEventCreate(&s.t1);
        }
        if (he != hipSuccess) {
            for (IoSlot &s : slots)
                for (hipEvent_t ev : {s.in_done, s.k_done, s.done, s.t0, s.t1})
                    if (ev) (void)hipEventDestroy(ev);
            for (hipStream_t x : st)
                if (x) (void)hipStreamDestroy(x);
            return hip_fail(he, "pipeline streams");
        }
        ctx->io_copy = st[0];
        ctx->io_kern = st[1];
        ctx->io = std::move(slots);
    }
    // chunks of about kChunkItems, none below kMinChunkItems unless the call is smaller
    size_t n_chunks = std::max<size_t>(1, (count + kChunkItems - 1) / kChunkItems);
    if (n_chunks > 1 && count / n_chunks < kMinChunkItems) n_chunks = std::max<size_t>(1, count / kMinChunkItems);
    const size_t per = (count + n_chunks - 1) / n_chunks;
    n_chunks = (count + per - 1) / per;
    const size_t kem_e_stride = (max_e + 3) / 4 * 4;
    const bool pinned_in = is_pinned(ctx, ct, count * ct_stride);
    const bool pinned_out = is_pinned(ctx, out, count * out_stride);

    int rc = CMD_OK;
    std::vector<FusedTile> tiles;
    std::vector<int32_t> items, item_key;

    // Host work and copy in of chunk c into slot c % kIoSlots: plan, staging, the copies of
    // the plan image and the ciphertexts, their event, and the kernel stream's wait for it.
    auto stage_in = [&](size_t c) -> int {
        const size_t first = c * per, n = std::min(per, count - first);
        IoSlot &s = ctx->io[c % kIoSlots];
        int r = io_finish(ctx, s, out, out_stride);
        if (r) return r;
        TraceRange tp("pipeline plan");
        fit_tiles(ctx, ctx->sched, n, key_ids + first, s.sched);
        r = plan_fused(ctx, s.sched, n, key_ids + first, s.tiles, items, item_key, &s.launch_tile);
        if (r) return r;
        // plan image: [items | item_key | tiles | counter]
        s.off_keys = (n * 4 + 255) / 256 * 256;
        s.off_tiles = s.off_keys + s.off_keys;
        s.off_ctr = s.off_tiles + (s.tiles.size() * sizeof(FusedTile) + 255) / 256 * 256;
        const size_t plan_bytes = s.off_ctr + 256;
        void *hp = s.h_plan;
        r = host_reserve(hp, s.h_plan_cap, plan_bytes);
        s.h_plan = (char *)hp;
        if (!r) r = dev_reserve(ctx, s.d_plan, plan_bytes);
        if (!r) r = dev_reserve(ctx, s.d_ct, n * ct_stride);
        if (!r) r = dev_reserve(ctx, s.d_out, n * out_stride);
        if (!r && (mode == CMD_MODE_KEM || mode == CMD_MODE_POSITIONS)) r = dev_reserve(ctx, s.d_kem_e, n * kem_e_stride);
        if (!r && mode == CMD_MODE_KEM) {
            r = dev_reserve(ctx, s.d_ok, n * 4);
            if (!r) r = dev_reserve(ctx, s.d_msg, n * (size_t)kKemMsgStride);
        }
        if (r) return r;
        std::memcpy(s.h_plan, items.data(), n * 4);
        std::memcpy(s.h_plan + s.off_keys, item_key.data(), n * 4);
        std::memcpy(s.h_plan + s.off_tiles, s.tiles.data(), s.tiles.size() * sizeof(FusedTile));
        const uint8_t *src = ct + first * ct_stride;
        if (!pinned_in) {
            void *hc = s.h_ct;
            r = host_reserve(hc, s.h_ct_cap, n * ct_stride);
            s.h_ct = (uint8_t *)hc;
            if (r) return r;
            copy_par(s.h_ct, src, n * ct_stride);
            src = s.h_ct;
        }
        s.staged_out = !pinned_out;
        s.dst = out + first * out_stride;
        if (s.staged_out) {
            void *ho = s.h_out;
            r = host_reserve(ho, s.h_out_cap, n * out_stride);
            s.h_out = (uint8_t *)ho;
            if (r) return r;
            s.dst = s.h_out;
        }
        s.first = first;
        s.count = n;
        s.out_bytes = n * out_stride;
        tp.end();
        hipError_t he = hipMemcpyAsync(s.d_plan.ptr, s.h_plan, s.off_ctr, hipMemcpyHostToDevice, ctx->io_copy);
        if (he == hipSuccess) he = hipMemcpyAsync(s.d_ct.ptr, src, n * ct_stride, hipMemcpyHostToDevice, ctx->io_copy);
        if (he == hipSuccess) he = hipEventRecord(s.in_done, ctx->io_copy);
        if (he == hipSuccess) he = hipStreamWaitEvent(ctx->io_kern, s.in_done, 0);
        return he == hipSuccess ? CMD_OK : hip_fail(he, "pipeline copy in");
    };

    // The kernels of a staged chunk on the kernel stream, then its `k_done` event.
    auto launch = [&](size_t c) -> int {
        IoSlot &s = ctx->io[c % kIoSlots];
        char *dp = (char *)s.d_plan.ptr;
        // rows come back whole: bytes past an item's output read as zero
        hipError_t he = hipMemsetAsync(s.d_out.ptr, 0, s.out_bytes, ctx->io_kern);
        s.timed = ctx->timing;
        if (he == hipSuccess && s.timed) he = hipEventRecord(s.t0, ctx->io_kern);
        if (he != hipSuccess) return hip_fail(he, "pipeline kernels");
        int r = enqueue_fused(ctx, mode, s.sched, s.tiles, s.launch_tile, (const FusedTile *)(dp + s.off_tiles),
                              (const int32_t *)dp, (const int32_t *)(dp + s.off_keys), (uint32_t *)(dp + s.off_ctr),
                              (const uint8_t *)s.d_ct.ptr, ct_stride, (uint8_t *)s.d_out.ptr, out_stride,
                              KemScratch{(uint8_t *)s.d_kem_e.ptr, kem_e_stride, (uint32_t *)s.d_ok.ptr,
                                         (uint8_t *)s.d_msg.ptr},
                              ctx->io_kern);
        if (r) return r;
        if (s.timed) he = hipEventRecord(s.t1, ctx->io_kern);
        if (he == hipSuccess) he = hipEventRecord(s.k_done, ctx->io_kern);
        if (he != hipSuccess) return hip_fail(he, "pipeline kernels");
        s.busy = true;
        s.out_pending = true;
        return CMD_OK;
    };

    // Per chunk c: chunk c-1's copy out (waits for its kernels), chunk c's kernels, chunk c+1's
    // host work and copy in, then the event marking chunk c-1's copy out done.  Errors end the
    // loop; what is in flight is drained below.
    size_t launched = 0;
    rc = stage_in(0);
    for (size_t c = 0; c < n_chunks && !rc; c++) {
        if (c > 0) rc = io_copy_out(ctx, ctx->io[(c - 1) % kIoSlots]);
        if (!rc) rc = launch(c);
        if (rc) break;
        launched = c + 1;
        if (c + 1 < n_chunks) rc = stage_in(c + 1);
        if (c > 0) {
            const int r2 = io_mark_done(ctx, ctx->io[(c - 1) % kIoSlots]);
            if (!rc) rc = r2;
        }
    }
    // the last launched chunk's copy out, then drain oldest first (also after an error, so no
    // copy is left running)
    if (launched > 0) {
        IoSlot &last = ctx->io[(launched - 1) % kIoSlots];
        int r2 = io_copy_out(ctx, last);
        if (!r2) r2 = io_mark_done(ctx, last);
        if (!rc) rc = r2;
    }
    for (size_t k = 0; k < kIoSlots; k++) {
        IoSlot &s = ctx->io[(launched + k) % kIoSlots];
        const int r2 = io_finish(ctx, s, out, out_stride);
        if (!rc) rc = r2;
    }
    if (rc) {
        // a chunk that failed part-way may still have copies or kernels queued
        (void)hipStreamSynchronize(ctx->io_kern);
        (void)hipStreamSynchronize(ctx->io_copy);
    }
    return rc;
}

}  // namespace cmd

extern "C" int cmd_host_alloc(cmd_ctx *ctx, size_t bytes, void **ptr)
{
    if (!ctx || !ptr || !bytes) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    ctx->host_pinned.emplace_back((uintptr_t)p, bytes);
    *ptr = p;
    return CMD_OK;
}

extern "C" int cmd_host_free(cmd_ctx *ctx, void *ptr)
{
    if (!ctx || !ptr) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    for (size_t i = 0; i < ctx->host_pinned.size(); i++) {
        if (ctx->host_pinned[i].first == (uintptr_t)ptr) {
            (void)hipHostFree(ptr);
            ctx->host_pinned.erase(ctx->host_pinned.begin() + (long)i);
            return CMD_OK;
        }
    }
    return set_error(CMD_ERR_ARG, "pointer was not allocated by cmd_host_alloc on this context");
}
