// Host-buffer decryption pipeline and pinned host memory.
//
// cmd_decrypt on host buffers (fused, KEM and positions modes) cuts the items into chunks that flow
// through kIoSlots slots.  Each slot owns a stream, device buffers for one chunk and pinned
// staging; a chunk's work list, ciphertexts, kernels and results are all enqueued on the
// slot's stream, so the copy in of one chunk, the kernels of another and the copy out of a
// third overlap (the reference's driver runs them back to back per BATCH_SIZE chunk,
// GPU_Baseline/param/param8192128/src/cuda/Decrypt.cu:177-324).  Buffers the caller got from
// cmd_host_alloc are already pinned: they are copied to and from the device directly;
// other (pageable) buffers go through the slot's staging with multi-threaded copies.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.h"

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

// Copies a finished slot's results to the caller (staged chunks only).
int io_finish(cmd_ctx *ctx, IoSlot &s, uint8_t *out, size_t out_stride)
{
    if (!s.busy) return CMD_OK;
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
    for (IoSlot &s : ctx->io) {
        if (s.st) (void)hipStreamSynchronize(s.st);
        for (DevBuf *b : {&s.d_ct, &s.d_out, &s.d_plan, &s.d_kem_e, &s.d_ok, &s.d_msg}) dev_free(ctx, *b);
        if (s.h_ct) (void)hipHostFree(s.h_ct);
        if (s.h_out) (void)hipHostFree(s.h_out);
        if (s.h_plan) (void)hipHostFree(s.h_plan);
        for (hipEvent_t ev : {s.done, s.t0, s.t1})
            if (ev) (void)hipEventDestroy(ev);
        if (s.st) (void)hipStreamDestroy(s.st);
    }
    ctx->io.clear();
    for (const auto &r : ctx->host_pinned) (void)hipHostFree((void *)r.first);
    ctx->host_pinned.clear();
}

int io_decrypt(cmd_ctx *ctx, int mode, size_t count, const int32_t *key_ids, const uint8_t *ct, size_t ct_stride,
               uint8_t *out, size_t out_stride)
{
    // validate every item first: a bad key must not leave half the output written
    size_t max_ct = 0, max_e = 0;
    for (size_t i = 0; i < count; i++) {
        const int32_t k = key_ids[i];
        if (k < 0 || (size_t)k >= ctx->keys.size() || !ctx->keys[k].used)
            return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)k);
        const ParamDesc &p = kParams[ctx->keys[k].pidx];
        max_ct = std::max(max_ct, (size_t)p.synd_bytes);
        max_e = std::max(max_e, (size_t)p.n / 8);
    }
    const size_t need_out = out_row_bytes(mode, count, key_ids, ctx);
    if (mode == CMD_MODE_POSITIONS && (out_stride & 1u))
        return set_error(CMD_ERR_ARG, "position rows need an even stride (%zu)", out_stride);
    if (ct_stride < max_ct || out_stride < need_out)
        return set_error(CMD_ERR_ARG, "strides (%zu, %zu) below item sizes (%zu, %zu)", ct_stride, out_stride, max_ct,
                         need_out);
    if (ctx->io.empty()) {
        // all slots or none: a failure leaves no half-made slot behind
        std::vector<IoSlot> slots(kIoSlots);
        hipError_t he = hipSuccess;
        for (IoSlot &s : slots) {
            if (he == hipSuccess) he = hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
            if (he == hipSuccess) he = hipEventCreate(&s.t0);
            if (he == hipSuccess) he = hipEventCreate(&s.t1);
        }
        if (he != hipSuccess) {
            for (IoSlot &s : slots) {
                for (hipEvent_t ev : {s.done, s.t0, s.t1})
                    if (ev) (void)hipEventDestroy(ev);
                if (s.st) (void)hipStreamDestroy(s.st);
            }
            return hip_fail(he, "pipeline streams");
        }
        ctx->io = std::move(slots);
    }
    // chunks of about kChunkItems, none below kMinChunkItems unless the call is smaller
    size_t n_chunks = std::max<size_t>(1, (count + kChunkItems - 1) / kChunkItems);
    if (n_chunks > 1 && count / n_chunks < kMinChunkItems) n_chunks = std::max<size_t>(1, count / kMinChunkItems);
    const size_t per = (count + n_chunks - 1) / n_chunks;
    const size_t kem_e_stride = (max_e + 3) / 4 * 4;
    const bool pinned_in = is_pinned(ctx, ct, count * ct_stride);
    const bool pinned_out = is_pinned(ctx, out, count * out_stride);

    int rc = CMD_OK;
    size_t slot_i = 0;
    std::vector<FusedTile> tiles;
    std::vector<int32_t> items, item_key;
    for (size_t first = 0; first < count && !rc; first += per, slot_i++) {
        const size_t n = std::min(per, count - first);
        IoSlot &s = ctx->io[slot_i % kIoSlots];
        rc = io_finish(ctx, s, out, out_stride);
        if (rc) break;
        int launch_tile = 32;
        FusedSchedule sched[kNumParams];
        fit_tiles(ctx, ctx->sched, n, key_ids + first, sched);
        rc = plan_fused(ctx, sched, n, key_ids + first, tiles, items, item_key, &launch_tile);
        if (rc) break;
        // plan image: [items | item_key | tiles | counter]
        const size_t off_keys = (n * 4 + 255) / 256 * 256;
        const size_t off_tiles = off_keys + off_keys;
        const size_t off_ctr = off_tiles + (tiles.size() * sizeof(FusedTile) + 255) / 256 * 256;
        const size_t plan_bytes = off_ctr + 256;
        void *hp = s.h_plan;
        rc = host_reserve(hp, s.h_plan_cap, plan_bytes);
        s.h_plan = (char *)hp;
        if (!rc) rc = dev_reserve(ctx, s.d_plan, plan_bytes);
        if (!rc) rc = dev_reserve(ctx, s.d_ct, n * ct_stride);
        if (!rc) rc = dev_reserve(ctx, s.d_out, n * out_stride);
        if (!rc && (mode == CMD_MODE_KEM || mode == CMD_MODE_POSITIONS)) rc = dev_reserve(ctx, s.d_kem_e, n * kem_e_stride);
        if (!rc && mode == CMD_MODE_KEM) {
            rc = dev_reserve(ctx, s.d_ok, n * 4);
            if (!rc) rc = dev_reserve(ctx, s.d_msg, n * (size_t)kKemMsgStride);
        }
        if (rc) break;
        std::memcpy(s.h_plan, items.data(), n * 4);
        std::memcpy(s.h_plan + off_keys, item_key.data(), n * 4);
        std::memcpy(s.h_plan + off_tiles, tiles.data(), tiles.size() * sizeof(FusedTile));
        const uint8_t *src = ct + first * ct_stride;
        if (!pinned_in) {
            void *hc = s.h_ct;
            rc = host_reserve(hc, s.h_ct_cap, n * ct_stride);
            s.h_ct = (uint8_t *)hc;
            if (rc) break;
            copy_par(s.h_ct, src, n * ct_stride);
            src = s.h_ct;
        }
        s.staged_out = !pinned_out;
        uint8_t *dst = out + first * out_stride;
        if (s.staged_out) {
            void *ho = s.h_out;
            rc = host_reserve(ho, s.h_out_cap, n * out_stride);
            s.h_out = (uint8_t *)ho;
            if (rc) break;
            dst = s.h_out;
        }
        // errors end the loop (no early return): the chunks already in flight are drained below
        char *dp = (char *)s.d_plan.ptr;
        hipError_t he = hipMemcpyAsync(dp, s.h_plan, off_ctr, hipMemcpyHostToDevice, s.st);
        if (he == hipSuccess) he = hipMemcpyAsync(s.d_ct.ptr, src, n * ct_stride, hipMemcpyHostToDevice, s.st);
        // rows come back whole: bytes past an item's output read as zero
        if (he == hipSuccess) he = hipMemsetAsync(s.d_out.ptr, 0, n * out_stride, s.st);
        s.timed = ctx->timing;
        if (he == hipSuccess && s.timed) he = hipEventRecord(s.t0, s.st);
        if (he != hipSuccess) {
            rc = hip_fail(he, "pipeline copy in");
            break;
        }
        rc = enqueue_fused(ctx, mode, sched, tiles, launch_tile, (const FusedTile *)(dp + off_tiles),
                           (const int32_t *)dp, (const int32_t *)(dp + off_keys), (uint32_t *)(dp + off_ctr),
                           (const uint8_t *)s.d_ct.ptr, ct_stride, (uint8_t *)s.d_out.ptr, out_stride,
                           KemScratch{(uint8_t *)s.d_kem_e.ptr, kem_e_stride, (uint32_t *)s.d_ok.ptr,
                                      (uint8_t *)s.d_msg.ptr},
                           s.st);
        if (rc) break;
        if (s.timed) he = hipEventRecord(s.t1, s.st);
        if (he == hipSuccess) he = hipMemcpyAsync(dst, s.d_out.ptr, n * out_stride, hipMemcpyDeviceToHost, s.st);
        if (he == hipSuccess) he = hipEventRecord(s.done, s.st);
        if (he != hipSuccess) {
            rc = hip_fail(he, "pipeline copy out");
            (void)hipStreamSynchronize(s.st);
            break;
        }
        s.busy = true;
        s.first = first;
        s.count = n;
    }
    // drain in chunk order (also after an error, so no copy is left running)
    for (size_t k = 0; k < kIoSlots; k++) {
        IoSlot &s = ctx->io[(slot_i + k) % kIoSlots];
        const int r2 = io_finish(ctx, s, out, out_stride);
        if (!rc) rc = r2;
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
