// Fused decryption on device buffers without any device allocation.
//
// cmd_decrypt_device plans each call into a work list (tiles, item order, per-item keys and
// the work-queue counter) that it allocates on the device.  Callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc) use
// cmd_decrypt_device_noalloc instead: the work list goes to `scratch`, a fixed array of the
// library in each device's memory, held until the process ends.  Host code never passes the
// array's name to the device: the address of each device's copy is looked up once per
// device (noalloc_setup, from cmd_ctx_create with that device current) and kept in
// g_scratch[device].  Every call on a device shares that device's array, so calls that
// overlap on one GPU race; calls through one context are serialised by its lock, and each
// call waits for its kernels before returning.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.h"

namespace cmd {

namespace {

constexpr size_t kNoallocItems = CMD_NOALLOC_MAX_ITEMS;
constexpr size_t kOffItems = 0;
constexpr size_t kOffKeys = kOffItems + kNoallocItems * 4;
constexpr size_t kOffTiles = kOffKeys + kNoallocItems * 4;
constexpr size_t kOffCounter = kOffTiles + kNoallocItems * sizeof(FusedTile);  // a tile holds >= 1 item
constexpr size_t kScratchBytes = kOffCounter + 256;
static_assert(kOffTiles % 16 == 0 && kOffCounter % 256 == 0, "aligned sections");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

// 6 MiB per device for CMD_NOALLOC_MAX_ITEMS items; host code reaches it only through
// g_scratch[device].
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return set_error(CMD_ERR_NODEV, "device %d", device);
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return CMD_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = p;
    return CMD_OK;
}

}  // namespace cmd

using namespace cmd;

extern "C" int cmd_decrypt_device_noalloc(cmd_ctx *ctx, size_t count, const int32_t *key_ids, const void *d_ct,
                                          size_t ct_stride, void *d_e, size_t e_stride, void *stream)
{
    if (!ctx) return CMD_ERR_ARG;
    if (!count) return CMD_OK;
    if (!key_ids || !d_ct || !d_e) return CMD_ERR_ARG;
    if (count > kNoallocItems)
        return set_error(CMD_ERR_ARG, "%zu items: the allocation-free entry point takes at most %zu per call", count,
                         kNoallocItems);
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    }
    if (!base) return set_error(CMD_ERR_HIP, "device scratch of device %d was not set up", ctx->device);
    size_t max_ct = 0, max_e = 0;
    for (size_t i = 0; i < count; i++) {
        const int32_t k = key_ids[i];
        if (k < 0 || (size_t)k >= ctx->keys.size() || !ctx->keys[k].used)
            return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)k);
        const ParamDesc &p = kParams[ctx->keys[k].pidx];
        max_ct = std::max(max_ct, (size_t)p.synd_bytes);
        max_e = std::max(max_e, (size_t)p.n / 8);
    }
    if (ct_stride < max_ct || e_stride < max_e)
        return set_error(CMD_ERR_ARG, "strides (%zu, %zu) below item sizes (%zu, %zu)", ct_stride, e_stride, max_ct,
                         max_e);
    FusedSchedule sched[kNumParams];
    if (ctx->fit_tiles)
        fit_tiles(ctx, ctx->sched, count, key_ids, sched);
    else
        for (int i = 0; i < kNumParams; i++) sched[i] = ctx->sched[i];
    // host plan in the context's reusable vectors (no device memory is involved)
    int launch_tile = 32;
    int rc = plan_fused(ctx, sched, count, key_ids, ctx->na_tiles, ctx->na_items, ctx->na_item_key, &launch_tile);
    if (rc) return rc;
    char *b = (char *)base;
    int32_t *d_items = (int32_t *)(b + kOffItems);
    int32_t *d_item_key = (int32_t *)(b + kOffKeys);
    FusedTile *d_tiles = (FusedTile *)(b + kOffTiles);
    uint32_t *d_counter = (uint32_t *)(b + kOffCounter);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    CMD_HIP(hipMemcpyAsync(d_items, ctx->na_items.data(), ctx->na_items.size() * 4, hipMemcpyHostToDevice, st));
    CMD_HIP(hipMemcpyAsync(d_item_key, ctx->na_item_key.data(), ctx->na_item_key.size() * 4, hipMemcpyHostToDevice,
                           st));
    CMD_HIP(hipMemcpyAsync(d_tiles, ctx->na_tiles.data(), ctx->na_tiles.size() * sizeof(FusedTile),
                           hipMemcpyHostToDevice, st));
    // the work-queue counter is zeroed by a kernel in the launch sequence (enqueue_fused)
    rc = enqueue_fused(ctx, CMD_MODE_FUSED, sched, ctx->na_tiles, launch_tile, d_tiles, d_items, d_item_key, d_counter,
                       (const uint8_t *)d_ct, ct_stride, (uint8_t *)d_e, e_stride, KemScratch{nullptr, 0, nullptr, nullptr},
                       st);
    // the shared array and the host plan must outlive this call's kernels and copies
    const hipError_t se = hipStreamSynchronize(st);
    if (rc) return rc;
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize");
    return CMD_OK;
}
