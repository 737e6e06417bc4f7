// Host engine behind the C ABI: contexts, resident keys, batch planning and launch.
#include "engine.h"
#include "trace.h"

#include <algorithm>
#include <thread>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace cmd {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(CMD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int dev_alloc(cmd_ctx *ctx, DevBuf &b, size_t bytes)
{
    dev_free(ctx, b);
    if (!bytes) return CMD_OK;
    if (ctx->used + bytes > ctx->budget)
        return set_error(CMD_ERR_BUDGET, "device budget: %llu used + %zu requested > %llu",
                         (unsigned long long)ctx->used, bytes, (unsigned long long)ctx->budget);
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    b.bytes = bytes;
    ctx->used += bytes;
    return CMD_OK;
}

int dev_reserve(cmd_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes && b.ptr) return CMD_OK;
    size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    int rc = dev_alloc(ctx, b, want);
    if (rc == CMD_ERR_BUDGET && want > bytes) rc = dev_alloc(ctx, b, bytes);
    return rc;
}

void dev_free(cmd_ctx *ctx, DevBuf &b)
{
    if (b.ptr) {
        (void)hipFree(b.ptr);
        ctx->used -= b.bytes;
    }
    b.ptr = nullptr;
    b.bytes = 0;
}

FusedSchedule default_schedule(int pidx)
{
    (void)pidx;
    FusedSchedule s;
    s.tile_cts = 32;
    s.waves_per_group = 4;
    s.waves_per_ct = 1;
    s.groups_per_cu = 2;
    s.root_alg = 1;
    return s;
}

// Offsets of the arrays inside a key's single device allocation, and its size.
struct KeyLayout {
    size_t off_L, off_lb, off_bn, off_y, off_s, off_g, bytes;
};

static KeyLayout key_layout(int pidx)
{
    const ParamDesc &p = kParams[pidx];
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t words = (size_t)1 << (p.m - 5);
    KeyLayout l;
    l.off_L = up((size_t)p.sb * 2 * p.t * 2);
    l.off_lb = l.off_L + up((size_t)p.n * 2);
    l.off_bn = l.off_lb + up((size_t)p.m * lbits_stride(pidx) * 4);
    l.off_y = l.off_bn + up((size_t)(2 * p.m - 1) * words * 4);
    l.off_s = l.off_y + up((size_t)p.m * words * 4);
    l.off_g = l.off_s + up((size_t)p.n / 8);
    l.bytes = l.off_g + up((size_t)2 * p.t);
    return l;
}

static size_t key_device_bytes(int pidx) { return key_layout(pidx).bytes; }

static int upload_key_slot(cmd_ctx *ctx, int32_t id)
{
    CMD_HIP(hipMemcpy(ctx->d_keys + id, &ctx->keys[id].dev, sizeof(KeyDev), hipMemcpyHostToDevice));
    return CMD_OK;
}

// Waits for the work of the context's own streams (its default stream and the host-buffer
// pipeline's copy and kernel streams); other streams of the device are left alone.
static void ctx_quiesce(cmd_ctx *ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t st : {ctx->io_copy, ctx->io_kern})
        if (st) (void)hipStreamSynchronize(st);
}

// ---- batch planning ---------------------------------------------------------------------

void batch_release(cmd_batch *b)
{
    for (int32_t k : b->pinned) b->ctx->keys[k].inflight--;
    b->pinned.clear();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    dev_free(b->ctx, b->d_plan);
    dev_free(b->ctx, b->d_s);
    dev_free(b->ctx, b->d_loc);
    dev_free(b->ctx, b->d_kem_e);
    dev_free(b->ctx, b->d_ok);
    dev_free(b->ctx, b->d_msg);
}

int plan_fused(const cmd_ctx *ctx, const FusedSchedule sched[], size_t count, const int32_t *key_ids,
               std::vector<FusedTile> &tiles, std::vector<int32_t> &items, std::vector<int32_t> &item_key,
               int *launch_tile)
{
    const size_t K = ctx->keys.size();
    for (size_t i = 0; i < count; i++) {
        const uint32_t k = (uint32_t)key_ids[i];
        if (k >= K || !ctx->keys[k].used || ctx->keys[k].pidx < 0)
            return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)key_ids[i]);
    }
    // items ordered by set, heavier sets (larger index here) first, then by key id, then by
    // position; set_end[p] is where set p's items end in that order
    items.resize(count);
    item_key.resize(count);
    size_t set_end[kNumParams] = {};
    if (count < K) {
        // few items against many key slots (serving launches): a stable comparison sort
        size_t per_set[kNumParams] = {};
        for (size_t i = 0; i < count; i++) per_set[ctx->keys[key_ids[i]].pidx]++;
        size_t pos = 0;
        for (int pidx = kNumParams - 1; pidx >= 0; pidx--) set_end[pidx] = pos += per_set[pidx];
        std::iota(items.begin(), items.end(), 0);
        std::stable_sort(items.begin(), items.end(), [&](int32_t a, int32_t c) {
            const int pa = ctx->keys[key_ids[a]].pidx, pc = ctx->keys[key_ids[c]].pidx;
            if (pa != pc) return pa > pc;
            return key_ids[a] < key_ids[c];
        });
    } else {
        // a counting sort over key ids, linear in the items and the key slots
        std::vector<size_t> at(K, 0);
        for (size_t i = 0; i < count; i++) at[(size_t)key_ids[i]]++;
        size_t pos = 0;
        for (int pidx = kNumParams - 1; pidx >= 0; pidx--) {
            for (size_t k = 0; k < K; k++)
                if (ctx->keys[k].used && ctx->keys[k].pidx == pidx) {
                    const size_t c = at[k];
                    at[k] = pos;
                    pos += c;
                }
            set_end[pidx] = pos;
        }
        for (size_t i = 0; i < count; i++) items[at[(size_t)key_ids[i]]++] = (int32_t)i;
    }
    for (size_t i = 0; i < count; i++) item_key[i] = key_ids[items[i]];
    tiles.clear();
    int widest = 0;
    // tiles of up to tile_cts consecutive items of one set (a tile may span keys of that set)
    size_t set_begin = 0;
    for (int pidx = kNumParams - 1; pidx >= 0; pidx--) {
        const size_t width = (size_t)sched[pidx].tile_cts;
        if (set_end[pidx] > set_begin) widest = std::max(widest, sched[pidx].tile_cts);
        for (size_t i = set_begin; i < set_end[pidx]; i += width) {
            FusedTile t;
            t.pidx = pidx;
            t.first = (int32_t)i;
            t.count = (int32_t)std::min(width, set_end[pidx] - i);
            t.pad = 0;
            tiles.push_back(t);
        }
        set_begin = set_end[pidx];
    }
    if (launch_tile) *launch_tile = widest ? widest : 32;
    return CMD_OK;
}

FusedSchedule launch_schedule(const FusedSchedule sched[], const std::vector<FusedTile> &tiles, int launch_tile)
{
    FusedSchedule s = tiles.empty() ? default_schedule(0) : sched[tiles[0].pidx];
    s.tile_cts = launch_tile;
    return s;
}

void fit_tiles(const cmd_ctx *ctx, const FusedSchedule in[], size_t count, const int32_t *key_ids,
               FusedSchedule out[])
{
    size_t per_set[kNumParams] = {};
    if (count < ctx->keys.size()) {
        for (size_t i = 0; i < count; i++) {
            const uint32_t k = (uint32_t)key_ids[i];
            if (k < ctx->keys.size() && ctx->keys[k].used && ctx->keys[k].pidx >= 0) per_set[ctx->keys[k].pidx]++;
        }
    } else {
        // many items: counted per key first, then per set
        std::vector<size_t> per_key(ctx->keys.size(), 0);
        for (size_t i = 0; i < count; i++) {
            const uint32_t k = (uint32_t)key_ids[i];
            if (k < per_key.size()) per_key[k]++;
        }
        for (size_t k = 0; k < per_key.size(); k++)
            if (per_key[k] && ctx->keys[k].used && ctx->keys[k].pidx >= 0) per_set[ctx->keys[k].pidx] += per_key[k];
    }
    for (int i = 0; i < kNumParams; i++) out[i] = in[i];
    const size_t target = (size_t)ctx->n_cu * 4;
    for (;;) {
        size_t tiles = 0;
        bool narrower = false;
        for (int i = 0; i < kNumParams; i++) {
            if (!per_set[i]) continue;
            tiles += (per_set[i] + out[i].tile_cts - 1) / out[i].tile_cts;
            narrower |= out[i].tile_cts > kMinFitTile;
        }
        if (tiles >= target || !narrower) break;
        for (int i = 0; i < kNumParams; i++)
            if (out[i].tile_cts > kMinFitTile) out[i].tile_cts >>= 1;
    }
}

size_t out_row_bytes(int mode, size_t count, const int32_t *key_ids, const cmd_ctx *ctx)
{
    size_t e = 0, t = 0;
    for (size_t i = 0; i < count; i++) {
        const ParamDesc &p = kParams[ctx->keys[key_ids[i]].pidx];
        e = std::max(e, (size_t)p.n / 8);
        t = std::max(t, (size_t)p.t);
    }
    if (mode == CMD_MODE_KEM) return CMD_KEM_KEY_BYTES;
    if (mode == CMD_MODE_POSITIONS) return 2 * t;
    return e;
}

int batch_plan(cmd_batch *b, const int32_t *key_ids)
{
    cmd_ctx *ctx = b->ctx;
    const size_t count = b->count;
    // validate keys, find max sizes
    size_t max_ct = 0, max_e = 0;
    for (size_t i = 0; i < count; i++) {
        int32_t k = key_ids[i];
        if (k < 0 || (size_t)k >= ctx->keys.size() || !ctx->keys[k].used)
            return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)k);
        const ParamDesc &p = kParams[ctx->keys[k].pidx];
        max_ct = std::max(max_ct, (size_t)p.synd_bytes);
        max_e = std::max(max_e, (size_t)p.n / 8);
    }
    const size_t max_out = out_row_bytes(b->mode, count, key_ids, ctx);
    if (b->ct_stride < max_ct || b->e_stride < max_out)
        return set_error(CMD_ERR_ARG, "strides (%zu, %zu) below item sizes (%zu, %zu)", b->ct_stride, b->e_stride,
                         max_ct, max_out);
    if (b->mode == CMD_MODE_POSITIONS && (b->e_stride & 1u))
        return set_error(CMD_ERR_ARG, "position rows need an even stride (%zu)", b->e_stride);

    // order items by (set, key), keeping input order inside a key
    std::vector<int32_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
        int pa = ctx->keys[key_ids[a]].pidx, pc = ctx->keys[key_ids[c]].pidx;
        if (pa != pc) return pa < pc;
        return key_ids[a] < key_ids[c];
    });

    b->groups.clear();
    b->tiles.clear();
    b->items.clear();
    b->item_key.clear();
    b->max_group = 0;
    b->max_t = 0;
    for (size_t i = 0; i < count;) {
        size_t j = i;
        while (j < count && key_ids[order[j]] == key_ids[order[i]]) j++;
        int key = key_ids[order[i]];
        int pidx = ctx->keys[key].pidx;
        b->groups.push_back({pidx, key, (int)i, (int)(j - i)});
        b->max_group = std::max(b->max_group, (int)(j - i));
        b->max_t = std::max(b->max_t, (int)kParams[pidx].t);
        i = j;
    }

    if (b->mode != CMD_MODE_PLAIN) {
        int rc = plan_fused(ctx, b->sched, count, key_ids, b->tiles, b->items, b->item_key, &b->launch_tile);
        if (rc) return rc;
    } else {
        b->items.assign(order.begin(), order.end());
    }

    // upload: [items | item_key | tiles | counter]
    size_t n_items = b->items.size();
    size_t off_items = 0;
    size_t off_keys = off_items + ((n_items * 4 + 255) / 256) * 256;
    size_t off_tiles = off_keys + ((b->item_key.size() * 4 + 255) / 256) * 256;
    size_t off_ctr = off_tiles + ((b->tiles.size() * sizeof(FusedTile) + 255) / 256) * 256;
    size_t total = off_ctr + 256;
    int rc = dev_alloc(ctx, b->d_plan, total);
    if (rc) return rc;
    char *base = (char *)b->d_plan.ptr;
    b->d_idx = (int32_t *)(base + off_items);
    b->d_items = (int32_t *)(base + off_items);
    b->d_item_key = (int32_t *)(base + off_keys);
    b->d_tiles = (FusedTile *)(base + off_tiles);
    b->d_counter = (uint32_t *)(base + off_ctr);
    if (n_items) CMD_HIP(hipMemcpy(b->d_items, b->items.data(), n_items * 4, hipMemcpyHostToDevice));
    if (!b->item_key.empty())
        CMD_HIP(hipMemcpy(b->d_item_key, b->item_key.data(), b->item_key.size() * 4, hipMemcpyHostToDevice));
    if (!b->tiles.empty())
        CMD_HIP(hipMemcpy(b->d_tiles, b->tiles.data(), b->tiles.size() * sizeof(FusedTile), hipMemcpyHostToDevice));

    if (b->mode == CMD_MODE_PLAIN && b->max_group > 0) {
        rc = dev_alloc(ctx, b->d_s, (size_t)b->max_group * 2 * b->max_t * 2);
        if (!rc) rc = dev_alloc(ctx, b->d_loc, (size_t)b->max_group * (b->max_t + 1) * 2);
        if (rc) return rc;
    }
    if ((b->mode == CMD_MODE_KEM || b->mode == CMD_MODE_POSITIONS) && count > 0) {
        // decoded error vectors (and for KEM accept flags and SHAKE256 inputs) stay on the device
        b->kem_e_stride = (max_e + 3) / 4 * 4;
        rc = dev_alloc(ctx, b->d_kem_e, count * b->kem_e_stride);
        if (!rc && b->mode == CMD_MODE_KEM) rc = dev_alloc(ctx, b->d_ok, count * 4);
        if (!rc && b->mode == CMD_MODE_KEM) rc = dev_alloc(ctx, b->d_msg, count * (size_t)kKemMsgStride);
        if (rc) return rc;
    }
    return CMD_OK;
}

int batch_enqueue(cmd_batch *b, const void *d_ct, void *d_e, hipStream_t st)
{
    cmd_ctx *ctx = b->ctx;
    const uint8_t *ct = (const uint8_t *)d_ct;
    uint8_t *e = (uint8_t *)d_e;
    if (b->mode == CMD_MODE_PLAIN) {
        for (const auto &g : b->groups) {
            const KeyDev &key = ctx->keys[g.key].dev;
            const int32_t *idx = b->d_idx + g.first;
            {
                TraceRange r("synd");
                CMD_HIP(launch_plain_synd(g.pidx, key, g.count, idx, ct, b->ct_stride, (uint16_t *)b->d_s.ptr, st));
            }
            {
                TraceRange r("bm");
                CMD_HIP(launch_plain_bm(g.pidx, g.count, (const uint16_t *)b->d_s.ptr, (uint16_t *)b->d_loc.ptr, st));
            }
            {
                TraceRange r("chien");
                CMD_HIP(launch_plain_chien(g.pidx, key, g.count, idx, (const uint16_t *)b->d_loc.ptr, e, b->e_stride,
                                           st));
            }
        }
        return CMD_OK;
    }
    return enqueue_fused(ctx, b->mode, b->sched, b->tiles, b->launch_tile, b->d_tiles, b->d_items, b->d_item_key,
                         b->d_counter, ct, b->ct_stride, e, b->e_stride,
                         KemScratch{(uint8_t *)b->d_kem_e.ptr, b->kem_e_stride, (uint32_t *)b->d_ok.ptr,
                                    (uint8_t *)b->d_msg.ptr},
                         st);
}

int enqueue_fused(cmd_ctx *ctx, int mode, const FusedSchedule sched[], const std::vector<FusedTile> &tiles,
                  int launch_tile, const FusedTile *d_tiles, const int32_t *d_items, const int32_t *d_item_key,
                  uint32_t *d_counter, const uint8_t *ct, size_t ct_stride, uint8_t *out, size_t out_stride,
                  const KemScratch &kem, hipStream_t st)
{
    if (tiles.empty()) return CMD_OK;
    const bool is_kem = mode == CMD_MODE_KEM, to_scratch = is_kem || mode == CMD_MODE_POSITIONS;
    size_t n_items = 0;
    for (const FusedTile &t : tiles) n_items += (size_t)t.count;
    FusedArgs a;
    a.tiles = d_tiles;
    a.n_tiles = (int32_t)tiles.size();
    a.pad0 = 0;
    a.items = d_items;
    a.item_key = d_item_key;
    a.keys = ctx->d_keys;
    a.ct = ct;
    a.e = to_scratch ? kem.e : out;
    a.ct_stride = ct_stride;
    a.e_stride = to_scratch ? kem.e_stride : out_stride;
    a.tile_counter = d_counter;
    a.fft = ctx->d_fft;
    const FusedSchedule s = launch_schedule(sched, tiles, launch_tile);
    {
        TraceRange r("fused synd-bm-root");
        CMD_HIP(launch_zero_u32(d_counter, 1, st));
        CMD_HIP(launch_fused(a, s, ctx->n_cu, st));
    }
    if (mode == CMD_MODE_POSITIONS) {
        TraceRange r("positions");
        PositionsArgs pa;
        pa.count = (int32_t)n_items;
        pa.pad0 = 0;
        pa.items = d_items;
        pa.item_key = d_item_key;
        pa.keys = ctx->d_keys;
        pa.e = kem.e;
        pa.e_stride = kem.e_stride;
        pa.out = out;
        pa.out_stride = out_stride;
        CMD_HIP(launch_positions(pa, st));
    }
    if (!is_kem) return CMD_OK;

    DecapArgs d;
    d.tiles = d_tiles;
    d.n_tiles = (int32_t)tiles.size();
    d.pad0 = 0;
    d.items = d_items;
    d.item_key = d_item_key;
    d.keys = ctx->d_keys;
    d.ct = ct;
    d.e = kem.e;
    d.ct_stride = ct_stride;
    d.e_stride = kem.e_stride;
    d.ok = kem.ok;
    d.s_out = nullptr;
    d.fft = ctx->d_fft;
    {
        TraceRange r("kem check");
        CMD_HIP(launch_decap_check(d, st));
    }
    TraceRange r("kem hash");
    KemHashArgs h;
    h.count = (int32_t)n_items;
    h.pad0 = 0;
    h.items = d_items;
    h.item_key = d_item_key;
    h.keys = ctx->d_keys;
    h.ct = ct;
    h.e = kem.e;
    h.ct_stride = ct_stride;
    h.e_stride = kem.e_stride;
    h.ok = kem.ok;
    h.msg = kem.msg;
    h.key_out = out;
    h.key_stride = out_stride;
    CMD_HIP(launch_kem_hash(h, st));
    return CMD_OK;
}

}  // namespace cmd

using namespace cmd;

// ---- C ABI ------------------------------------------------------------------------------

extern "C" const char *cmd_strerror(int code)
{
    switch (code) {
    case CMD_OK: return "ok";
    case CMD_ERR_FILE: return "file open/read/write error";
    case CMD_ERR_CRYPTO: return "invalid count or key bytes";
    case CMD_ERR_PARAM: return "unknown parameter set";
    case CMD_ERR_ARG: return "invalid argument";
    case CMD_ERR_HIP: return "HIP runtime error";
    case CMD_ERR_NODEV: return "no GPU available";
    case CMD_ERR_BUDGET: return "device memory budget exceeded";
    case CMD_ERR_KEY: return "unknown key";
    case CMD_ERR_FULL: return "queue full";
    case CMD_ERR_NOMEM: return "out of host memory";
    case CMD_ERR_SINGULAR: return "no systematic form";
    case CMD_ERR_PENDING: return "request pending";
    case CMD_ERR_TIMEOUT: return "timed out";
    }
    return "unknown error";
}

extern "C" const char *cmd_last_error(void) { return g_last_error.c_str(); }

extern "C" int cmd_param_get(int32_t param_id, cmd_param_info *out)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (!out) return CMD_ERR_ARG;
    const ParamDesc &p = kParams[pidx];
    out->id = p.id;
    out->m = p.m;
    out->n = p.n;
    out->t = p.t;
    out->ct_bytes = p.synd_bytes;
    out->sb = p.sb;
    out->sk_bytes = p.sk_bytes;
    out->cond_bytes = p.cond_bytes;
    out->e_bytes = p.n / 8;
    return CMD_OK;
}

extern "C" int cmd_support_gen(int32_t param_id, const uint8_t *cond, uint16_t *L)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (!cond || !L) return CMD_ERR_ARG;
    return support_gen(pidx, cond, L);
}

extern "C" int cmd_key_expand_host(int32_t param_id, const uint8_t *sk, size_t sk_len, uint16_t *g, uint16_t *L,
                                   uint16_t *table)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    ExpandedKey K;
    int rc = expand_key(pidx, sk, sk_len, K);
    if (rc) return set_error(rc, "secret key of %zu bytes, %d expected", sk_len, kParams[pidx].sk_bytes);
    if (g) std::memcpy(g, K.g.data(), K.g.size() * 2);
    if (L) std::memcpy(L, K.L.data(), K.L.size() * 2);
    if (table) std::memcpy(table, K.table.data(), K.table.size() * 2);
    return CMD_OK;
}

// ---- errorstream text (write_error_positions, CPU/param/param6960119/src/decrypt.c:55-66) ----
// " %d" for every set bit in ascending order, then "\n", per vector.  Vectors are formatted in
// parallel chunks: each chunk first measures its text, then writes at its offset.

namespace {

struct DigitTable {
    char text[8192][4];  // decimal digits of 0 .. 8191, left aligned
    uint8_t len[8192];
    DigitTable()
    {
        for (int i = 0; i < 8192; i++) {
            char tmp[8];
            int n = 0, v = i;
            do { tmp[n++] = (char)('0' + v % 10); v /= 10; } while (v);
            len[i] = (uint8_t)n;
            for (int k = 0; k < n; k++) text[i][k] = tmp[n - 1 - k];
        }
    }
};

const DigitTable &digits()
{
    static const DigitTable t;
    return t;
}

// Bytes of one vector's line.
size_t line_bytes(const uint8_t *v, int nbytes)
{
    const DigitTable &d = digits();
    size_t pos = 1;  // "\n"
    int byte = 0;
    for (; byte + 8 <= nbytes; byte += 8) {
        uint64_t w;
        std::memcpy(&w, v + byte, 8);
        while (w) {
            pos += 1 + d.len[byte * 8 + __builtin_ctzll(w)];
            w &= w - 1;
        }
    }
    for (; byte < nbytes; byte++)
        for (uint32_t b = v[byte]; b; b &= b - 1) pos += 1 + d.len[byte * 8 + __builtin_ctz(b)];
    return pos;
}

// Writes one vector's line at out; returns its length.
size_t line_write(const uint8_t *v, int nbytes, char *out)
{
    const DigitTable &d = digits();
    char *p = out;
    auto put = [&](int i) {
        *p++ = ' ';
        std::memcpy(p, d.text[i], 4);  // the table row has room for 4 digits
        p += d.len[i];
    };
    int byte = 0;
    for (; byte + 8 <= nbytes; byte += 8) {
        uint64_t w;
        std::memcpy(&w, v + byte, 8);
        while (w) {
            put(byte * 8 + __builtin_ctzll(w));
            w &= w - 1;
        }
    }
    for (; byte < nbytes; byte++)
        for (uint32_t b = v[byte]; b; b &= b - 1) put(byte * 8 + __builtin_ctz(b));
    *p++ = '\n';
    return (size_t)(p - out);
}

// Runs fn(first, last) over [0, count) in up to hardware_concurrency chunks.
template <typename F>
void parallel_chunks(size_t count, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (count < 2048) nt = 1;
    std::vector<std::thread> pool;
    const size_t per = (count + nt - 1) / nt;
    for (unsigned t = 1; t < nt; t++) {
        const size_t a = t * per, b = std::min(count, a + per);
        if (a < b) pool.emplace_back(fn, a, b);
    }
    fn(0, std::min(count, per));
    for (auto &th : pool) th.join();
}

// Lines k < count measured by bytes_of(k) and written by write_line(k, dst) in parallel chunks.
template <typename Bytes, typename Write>
int format_lines(size_t count, Bytes bytes_of, Write write_line, char *buf, size_t cap, size_t *written)
{
    std::vector<size_t> off(count + 1, 0);
    parallel_chunks(count, [&](size_t a, size_t b) {
        for (size_t k = a; k < b; k++) off[k + 1] = bytes_of(k);
    });
    for (size_t k = 0; k < count; k++) off[k + 1] += off[k];
    const size_t total = off[count];
    if (written) *written = total;
    if (!buf || cap == 0) return CMD_OK;
    if (total > cap) return set_error(CMD_ERR_ARG, "buffer of %zu bytes, %zu needed", cap, total);
    // the last 3 bytes of a line may be overwritten by the 4-byte digit copy of the next
    // number; lines are written in order within a chunk, and each chunk's last line is
    // written with exact-size copies into a bounce buffer
    parallel_chunks(count, [&](size_t a, size_t b) {
        for (size_t k = a; k < b; k++) {
            const size_t len = off[k + 1] - off[k];
            if (k + 1 < b && off[k + 1] + 3 < total) {
                write_line(k, buf + off[k]);
            } else {
                std::vector<char> tmp(len + 4);
                write_line(k, tmp.data());
                std::memcpy(buf + off[k], tmp.data(), len);
            }
        }
    });
    return CMD_OK;
}

// Position rows: t entries, ascending, 0xFFFF (or any value >= n) ends the row.
size_t pos_line_bytes(const uint16_t *pos, int t, int n)
{
    const DigitTable &d = digits();
    size_t len = 1;
    for (int i = 0; i < t && pos[i] < n; i++) len += 1 + d.len[pos[i]];
    return len;
}

size_t pos_line_write(const uint16_t *pos, int t, int n, char *out)
{
    const DigitTable &d = digits();
    char *p = out;
    for (int i = 0; i < t && pos[i] < n; i++) {
        *p++ = ' ';
        std::memcpy(p, d.text[pos[i]], 4);
        p += d.len[pos[i]];
    }
    *p++ = '\n';
    return (size_t)(p - out);
}

}  // namespace

extern "C" int cmd_errorstream_format(int32_t param_id, const uint8_t *e, size_t e_stride, size_t count, char *buf,
                                      size_t cap, size_t *written)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    const int nbytes = kParams[pidx].n / 8;
    if ((!e && count) || e_stride < (size_t)nbytes) return CMD_ERR_ARG;
    return format_lines(
        count, [&](size_t k) { return line_bytes(e + k * e_stride, nbytes); },
        [&](size_t k, char *dst) { line_write(e + k * e_stride, nbytes, dst); }, buf, cap, written);
}

extern "C" int cmd_errorstream_format_positions(int32_t param_id, const uint16_t *pos, size_t pos_stride,
                                                size_t count, char *buf, size_t cap, size_t *written)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    const int t = kParams[pidx].t, n = kParams[pidx].n;
    if ((!pos && count) || pos_stride < (size_t)(2 * t) || (pos_stride & 1u)) return CMD_ERR_ARG;
    const size_t st = pos_stride / 2;
    return format_lines(
        count, [&](size_t k) { return pos_line_bytes(pos + k * st, t, n); },
        [&](size_t k, char *dst) { pos_line_write(pos + k * st, t, n, dst); }, buf, cap, written);
}

static int write_text(const char *path, const std::vector<char> &buf, int append)
{
    FILE *f = std::fopen(path, append ? "ab" : "wb");
    if (!f) return set_error(CMD_ERR_FILE, "cannot open %s", path);
    size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    int cl = std::fclose(f);
    if (w != buf.size() || cl != 0) return set_error(CMD_ERR_FILE, "short write to %s", path);
    return CMD_OK;
}

extern "C" int cmd_errorstream_write(int32_t param_id, const char *path, const uint8_t *e, size_t e_stride,
                                     size_t count, int append)
{
    size_t need = 0;
    int rc = cmd_errorstream_format(param_id, e, e_stride, count, nullptr, 0, &need);
    if (rc) return rc;
    std::vector<char> buf(need);
    if (need) rc = cmd_errorstream_format(param_id, e, e_stride, count, buf.data(), buf.size(), &need);
    return rc ? rc : write_text(path, buf, append);
}

extern "C" int cmd_errorstream_write_positions(int32_t param_id, const char *path, const uint16_t *pos,
                                               size_t pos_stride, size_t count, int append)
{
    size_t need = 0;
    int rc = cmd_errorstream_format_positions(param_id, pos, pos_stride, count, nullptr, 0, &need);
    if (rc) return rc;
    std::vector<char> buf(need);
    if (need) rc = cmd_errorstream_format_positions(param_id, pos, pos_stride, count, buf.data(), buf.size(), &need);
    return rc ? rc : write_text(path, buf, append);
}

extern "C" int cmd_ctx_create(int32_t device, uint64_t hbm_budget_bytes, cmd_ctx **out)
{
    if (!out) return CMD_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(CMD_ERR_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return set_error(CMD_ERR_NODEV, "device %d of %d", (int)device, ndev);
    CMD_HIP(hipSetDevice(device));
    if (int rc = noalloc_setup(device)) return rc;
    hipDeviceProp_t prop;
    CMD_HIP(hipGetDeviceProperties(&prop, device));
    std::unique_ptr<cmd_ctx> ctx(new cmd_ctx());
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    size_t free_b = 0, total_b = 0;
    CMD_HIP(hipMemGetInfo(&free_b, &total_b));
    ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(free_b * 0.9);
    CMD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    CMD_HIP(hipEventCreate(&ctx->ev0));
    CMD_HIP(hipEventCreate(&ctx->ev1));
    for (int i = 0; i < kNumParams; i++) ctx->sched[i] = default_schedule(i);
    CMD_HIP(hipMalloc(&ctx->d_keys, kMaxKeys * sizeof(KeyDev)));
    // unused slots read as null descriptors (pidx 0 with null tables), never as stale pointers
    CMD_HIP(hipMemset(ctx->d_keys, 0, kMaxKeys * sizeof(KeyDev)));
    {
        std::vector<FftTab> tabs(kFftVariants);
        build_fft_tab(12, 7, &tabs[0]);
        build_fft_tab(13, 7, &tabs[1]);
        build_fft_tab(13, 8, &tabs[2]);
        CMD_HIP(hipMalloc(&ctx->d_fft, kFftVariants * sizeof(FftTab)));
        CMD_HIP(hipMemcpy(ctx->d_fft, tabs.data(), kFftVariants * sizeof(FftTab), hipMemcpyHostToDevice));
    }
    *out = ctx.release();
    return CMD_OK;
}

extern "C" int cmd_ctx_destroy(cmd_ctx *ctx)
{
    if (!ctx) return CMD_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    io_release(ctx);
    for (auto &k : ctx->keys)
        if (k.mem) (void)hipFree(k.mem);
    if (ctx->d_keys) (void)hipFree(ctx->d_keys);
    if (ctx->d_fft) (void)hipFree(ctx->d_fft);
    for (DevBuf *b : {&ctx->ct, &ctx->e, &ctx->s, &ctx->loc, &ctx->idx, &ctx->plan, &ctx->keystage}) dev_free(ctx, *b);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CMD_OK;
}

extern "C" int cmd_ctx_memory(cmd_ctx *ctx, uint64_t *used, uint64_t *budget)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return CMD_OK;
}

extern "C" int cmd_ctx_set_budget(cmd_ctx *ctx, uint64_t hbm_budget_bytes)
{
    if (!ctx || !hbm_budget_bytes) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    ctx->budget = hbm_budget_bytes;
    return CMD_OK;
}

extern "C" int cmd_key_device_bytes(int32_t param_id, uint64_t *bytes)
{
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (bytes) *bytes = key_device_bytes(pidx);
    return CMD_OK;
}

extern "C" int cmd_key_load(cmd_ctx *ctx, int32_t param_id, const uint8_t *sk, size_t sk_len, int32_t *key_id)
{
    if (!ctx || !sk || !key_id) return CMD_ERR_ARG;
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    const ParamDesc &p = kParams[pidx];
    if (sk_len < (size_t)p.sk_bytes)
        return set_error(CMD_ERR_CRYPTO, "secret key of %zu bytes, %d expected", sk_len, p.sk_bytes);

    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    TraceRange tr("key load");
    CMD_HIP(hipSetDevice(ctx->device));
    // the Goppa coefficients, control bits and s go to the device, which expands them
    const size_t key_bytes = (size_t)2 * p.t + (size_t)p.cond_bytes + (size_t)p.n / 8;
    int rc = dev_reserve(ctx, ctx->keystage, key_bytes);
    if (rc) return rc;
    const size_t bytes = key_device_bytes(pidx);
    if (ctx->used + bytes > ctx->budget)
        return set_error(CMD_ERR_BUDGET, "key of %zu bytes exceeds the budget (%llu of %llu used)", bytes,
                         (unsigned long long)ctx->used, (unsigned long long)ctx->budget);
    void *mem = nullptr;
    CMD_HIP(hipMalloc(&mem, bytes));
    const KeyLayout lay = key_layout(pidx);
    char *base = (char *)mem;
    KeySlot slot;
    slot.used = true;
    slot.pidx = pidx;
    slot.mem = mem;
    slot.bytes = bytes;
    slot.dev.table = (const uint16_t *)base;
    slot.dev.L = (const uint16_t *)(base + lay.off_L);
    slot.dev.lbits = (const uint32_t *)(base + lay.off_lb);
    slot.dev.benes = (const uint32_t *)(base + lay.off_bn);
    slot.dev.ypl = (const uint32_t *)(base + lay.off_y);
    slot.dev.s = (const uint8_t *)(base + lay.off_s);
    slot.dev.gbytes = (const uint8_t *)(base + lay.off_g);
    slot.dev.pidx = pidx;
    slot.dev.lstride = lbits_stride(pidx);
    hipError_t e = hipMemcpyAsync(ctx->keystage.ptr, sk + 40, key_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_key_expand(pidx, (const uint8_t *)ctx->keystage.ptr, slot.dev, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(base + lay.off_s, (const char *)ctx->keystage.ptr + 2 * p.t + p.cond_bytes, (size_t)p.n / 8,
                           hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(base + lay.off_g, ctx->keystage.ptr, (size_t)2 * p.t, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(mem);
        return hip_fail(e, "key expansion");
    }
    int32_t id = -1;
    for (size_t i = 0; i < ctx->keys.size(); i++)
        if (!ctx->keys[i].used) { id = (int32_t)i; break; }
    if (id < 0) {
        if (ctx->keys.size() >= kMaxKeys) {
            (void)hipFree(mem);
            return set_error(CMD_ERR_BUDGET, "all %zu key slots are in use", kMaxKeys);
        }
        id = (int32_t)ctx->keys.size();
        ctx->keys.push_back(slot);
    } else {
        ctx->keys[id] = slot;
    }
    ctx->used += bytes;
    rc = upload_key_slot(ctx, id);
    if (rc) return rc;
    *key_id = id;
    return CMD_OK;
}

extern "C" int cmd_key_export(cmd_ctx *ctx, int32_t key_id, uint16_t *table, uint16_t *L, uint32_t *lbits,
                              uint32_t *benes, int32_t *lbits_stride_out)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    if (key_id < 0 || (size_t)key_id >= ctx->keys.size() || !ctx->keys[key_id].used)
        return set_error(CMD_ERR_KEY, "key id %d is not loaded", (int)key_id);
    CMD_HIP(hipSetDevice(ctx->device));
    const KeySlot &k = ctx->keys[key_id];
    const ParamDesc &p = kParams[k.pidx];
    if (lbits_stride_out) *lbits_stride_out = k.dev.lstride;
    if (table) CMD_HIP(hipMemcpy(table, k.dev.table, (size_t)p.sb * 2 * p.t * 2, hipMemcpyDeviceToHost));
    if (L) CMD_HIP(hipMemcpy(L, k.dev.L, (size_t)p.n * 2, hipMemcpyDeviceToHost));
    if (lbits) CMD_HIP(hipMemcpy(lbits, k.dev.lbits, (size_t)p.m * k.dev.lstride * 4, hipMemcpyDeviceToHost));
    if (benes)
        CMD_HIP(hipMemcpy(benes, k.dev.benes, (size_t)(2 * p.m - 1) * ((size_t)1 << (p.m - 5)) * 4,
                          hipMemcpyDeviceToHost));
    return CMD_OK;
}

extern "C" int cmd_key_unload(cmd_ctx *ctx, int32_t key_id)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    if (key_id < 0 || (size_t)key_id >= ctx->keys.size() || !ctx->keys[key_id].used)
        return set_error(CMD_ERR_KEY, "key id %d is not loaded", (int)key_id);
    if (ctx->keys[key_id].inflight > 0)
        return set_error(CMD_ERR_KEY, "key id %d has %d requests in flight", (int)key_id, ctx->keys[key_id].inflight);
    (void)hipSetDevice(ctx->device);
    // No batch or server request holds the key (inflight == 0) and every other entry point
    // waits for its own work; finish what the context's own streams may still run.
    ctx_quiesce(ctx);
    KeySlot &k = ctx->keys[key_id];
    (void)hipFree(k.mem);
    ctx->used -= k.bytes;
    k = KeySlot();
    return upload_key_slot(ctx, key_id);
}

static int make_batch(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids, size_t ct_stride,
                      size_t e_stride, int32_t use_graph, cmd_batch **out)
{
    if (mode != CMD_MODE_PLAIN && mode != CMD_MODE_FUSED && mode != CMD_MODE_KEM && mode != CMD_MODE_POSITIONS)
        return set_error(CMD_ERR_ARG, "mode %d", (int)mode);
    if (count && !key_ids) return CMD_ERR_ARG;
    if (count > (size_t)INT32_MAX) return set_error(CMD_ERR_ARG, "count %zu too large", count);
    std::unique_ptr<cmd_batch> b(new cmd_batch());
    b->ctx = ctx;
    b->mode = mode;
    b->count = count;
    b->ct_stride = ct_stride;
    b->e_stride = e_stride;
    b->use_graph = use_graph != 0;
    if (ctx->fit_tiles && mode != CMD_MODE_PLAIN && count)
        fit_tiles(ctx, ctx->sched, count, key_ids, b->sched);
    else
        for (int i = 0; i < kNumParams; i++) b->sched[i] = ctx->sched[i];
    int rc = batch_plan(b.get(), key_ids);
    if (!rc && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        b->done = nullptr;
        rc = set_error(CMD_ERR_HIP, "batch completion event");
    }
    if (rc) {
        batch_release(b.get());
        if (b->done) (void)hipEventDestroy(b->done);
        return rc;
    }
    // a captured graph holds the keys' device pointers: keep them loaded while the batch lives
    for (const auto &grp : b->groups) {
        ctx->keys[grp.key].inflight++;
        b->pinned.push_back(grp.key);
    }
    *out = b.release();
    return CMD_OK;
}

static int run_batch(cmd_batch *b, const void *d_ct, void *d_e, hipStream_t st)
{
    cmd_ctx *ctx = b->ctx;
    TraceRange tr(b->use_graph ? "batch graph" : "batch");
    if (ctx->timing) CMD_HIP(hipEventRecord(ctx->ev0, st));
    if (b->use_graph) {
        if (!b->exec || b->graph_ct != d_ct || b->graph_e != d_e || b->graph_stream != st) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
            CMD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = batch_enqueue(b, d_ct, d_e, st);
            hipGraph_t graph = nullptr;
            hipError_t ec = hipStreamEndCapture(st, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
            b->graph = graph;
            CMD_HIP(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
            b->graph_ct = d_ct;
            b->graph_e = d_e;
            b->graph_stream = st;
        }
        CMD_HIP(hipGraphLaunch(b->exec, st));
    } else {
        int rc = batch_enqueue(b, d_ct, d_e, st);
        if (rc) return rc;
    }
    if (b->done) {
        CMD_HIP(hipEventRecord(b->done, st));
        b->ran = true;
    }
    if (ctx->timing) {
        CMD_HIP(hipEventRecord(ctx->ev1, st));
        CMD_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        CMD_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->timing_us += ms * 1000.0;
        ctx->timing_n++;
    }
    return CMD_OK;
}

extern "C" int cmd_batch_create(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids, size_t ct_stride,
                                size_t e_stride, int32_t use_graph, cmd_batch **out)
{
    if (!ctx || !out) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    return make_batch(ctx, mode, count, key_ids, ct_stride, e_stride, use_graph, out);
}

extern "C" int cmd_batch_run(cmd_batch *b, const void *d_ct, void *d_e, void *stream)
{
    if (!b || (b->count && (!d_ct || !d_e))) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
    CMD_HIP(hipSetDevice(b->ctx->device));
    return run_batch(b, d_ct, d_e, stream ? (hipStream_t)stream : b->ctx->stream);
}

extern "C" int cmd_batch_destroy(cmd_batch *b)
{
    if (!b) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(b->ctx->lock);
    (void)hipSetDevice(b->ctx->device);
    // the last run's work (on whatever stream it was enqueued) must be done with the plan
    if (b->done && b->ran) (void)hipEventSynchronize(b->done);
    batch_release(b);
    if (b->done) (void)hipEventDestroy(b->done);
    delete b;
    return CMD_OK;
}

extern "C" int cmd_decrypt_device(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids, const void *d_ct,
                                  size_t ct_stride, void *d_e, size_t e_stride, void *stream)
{
    if (!ctx) return CMD_ERR_ARG;
    if (!count) return CMD_OK;
    if (!d_ct || !d_e) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    cmd_batch *b = nullptr;
    int rc = make_batch(ctx, mode, count, key_ids, ct_stride, e_stride, 0, &b);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    rc = run_batch(b, d_ct, d_e, st);
    // the plan buffers must outlive the launches
    hipError_t se = hipStreamSynchronize(st);
    batch_release(b);
    delete b;
    if (rc) return rc;
    if (se != hipSuccess) return hip_fail(se, "hipStreamSynchronize");
    return CMD_OK;
}

extern "C" int cmd_decrypt(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids, const uint8_t *ct,
                           size_t ct_stride, uint8_t *e_out, size_t e_stride)
{
    if (!ctx) return CMD_ERR_ARG;
    if (!count) return CMD_OK;
    if (!ct || !e_out || !key_ids) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    TraceRange tr("decrypt host buffers");
    CMD_HIP(hipSetDevice(ctx->device));
    if (mode == CMD_MODE_FUSED || mode == CMD_MODE_KEM || mode == CMD_MODE_POSITIONS)
        return io_decrypt(ctx, mode, count, key_ids, ct, ct_stride, e_out, e_stride);
    if (mode != CMD_MODE_PLAIN) return set_error(CMD_ERR_ARG, "mode %d", (int)mode);
    int rc = dev_reserve(ctx, ctx->ct, count * ct_stride);
    if (!rc) rc = dev_reserve(ctx, ctx->e, count * e_stride);
    if (rc) return rc;
    CMD_HIP(hipMemcpyAsync(ctx->ct.ptr, ct, count * ct_stride, hipMemcpyHostToDevice, ctx->stream));
    // rows come back whole: bytes past an item's n/8 read as zero rather than stale data
    CMD_HIP(hipMemsetAsync(ctx->e.ptr, 0, count * e_stride, ctx->stream));
    rc = cmd_decrypt_device(ctx, mode, count, key_ids, ctx->ct.ptr, ct_stride, ctx->e.ptr, e_stride, ctx->stream);
    if (rc) return rc;
    CMD_HIP(hipMemcpyAsync(e_out, ctx->e.ptr, count * e_stride, hipMemcpyDeviceToHost, ctx->stream));
    CMD_HIP(hipStreamSynchronize(ctx->stream));
    return CMD_OK;
}

extern "C" int cmd_kem_dec(cmd_ctx *ctx, size_t count, const int32_t *key_ids, const uint8_t *ct, size_t ct_stride,
                           uint8_t *key_out, size_t key_stride)
{
    return cmd_decrypt(ctx, CMD_MODE_KEM, count, key_ids, ct, ct_stride, key_out, key_stride);
}

// ---- per-stage entry points ----------------------------------------------------------------

static int check_key(cmd_ctx *ctx, int32_t key_id)
{
    if (key_id < 0 || (size_t)key_id >= ctx->keys.size() || !ctx->keys[key_id].used)
        return set_error(CMD_ERR_KEY, "key id %d is not loaded", (int)key_id);
    return CMD_OK;
}

extern "C" int cmd_stage_synd(cmd_ctx *ctx, int32_t key_id, size_t count, const uint8_t *ct, size_t ct_stride,
                              uint16_t *s_out)
{
    if (!ctx || (count && (!ct || !s_out))) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    int rc = check_key(ctx, key_id);
    if (rc) return rc;
    if (!count) return CMD_OK;
    CMD_HIP(hipSetDevice(ctx->device));
    const KeySlot &k = ctx->keys[key_id];
    const ParamDesc &p = kParams[k.pidx];
    if (ct_stride < (size_t)p.synd_bytes) return CMD_ERR_ARG;
    rc = dev_reserve(ctx, ctx->ct, count * ct_stride);
    if (!rc) rc = dev_reserve(ctx, ctx->s, count * 2 * p.t * 2);
    if (rc) return rc;
    CMD_HIP(hipMemcpyAsync(ctx->ct.ptr, ct, count * ct_stride, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(launch_plain_synd(k.pidx, k.dev, (int)count, nullptr, (const uint8_t *)ctx->ct.ptr, ct_stride,
                              (uint16_t *)ctx->s.ptr, ctx->stream));
    CMD_HIP(hipMemcpyAsync(s_out, ctx->s.ptr, count * 2 * p.t * 2, hipMemcpyDeviceToHost, ctx->stream));
    CMD_HIP(hipStreamSynchronize(ctx->stream));
    return CMD_OK;
}

extern "C" int cmd_stage_bm(cmd_ctx *ctx, int32_t param_id, size_t count, const uint16_t *s, uint16_t *loc_out)
{
    if (!ctx || (count && (!s || !loc_out))) return CMD_ERR_ARG;
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (!count) return CMD_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    const ParamDesc &p = kParams[pidx];
    int rc = dev_reserve(ctx, ctx->s, count * 2 * p.t * 2);
    if (!rc) rc = dev_reserve(ctx, ctx->loc, count * (p.t + 1) * 2);
    if (rc) return rc;
    CMD_HIP(hipMemcpyAsync(ctx->s.ptr, s, count * 2 * p.t * 2, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(launch_plain_bm(pidx, (int)count, (const uint16_t *)ctx->s.ptr, (uint16_t *)ctx->loc.ptr, ctx->stream));
    CMD_HIP(hipMemcpyAsync(loc_out, ctx->loc.ptr, count * (p.t + 1) * 2, hipMemcpyDeviceToHost, ctx->stream));
    CMD_HIP(hipStreamSynchronize(ctx->stream));
    return CMD_OK;
}

extern "C" int cmd_stage_root(cmd_ctx *ctx, int32_t key_id, size_t count, const uint16_t *loc, uint8_t *e_out,
                              size_t e_stride)
{
    if (!ctx || (count && (!loc || !e_out))) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    int rc = check_key(ctx, key_id);
    if (rc) return rc;
    if (!count) return CMD_OK;
    CMD_HIP(hipSetDevice(ctx->device));
    const KeySlot &k = ctx->keys[key_id];
    const ParamDesc &p = kParams[k.pidx];
    if (e_stride < (size_t)p.n / 8) return CMD_ERR_ARG;
    rc = dev_reserve(ctx, ctx->loc, count * (p.t + 1) * 2);
    if (!rc) rc = dev_reserve(ctx, ctx->e, count * e_stride);
    if (rc) return rc;
    CMD_HIP(hipMemcpyAsync(ctx->loc.ptr, loc, count * (p.t + 1) * 2, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(launch_plain_chien(k.pidx, k.dev, (int)count, nullptr, (const uint16_t *)ctx->loc.ptr,
                               (uint8_t *)ctx->e.ptr, e_stride, ctx->stream));
    CMD_HIP(hipMemcpyAsync(e_out, ctx->e.ptr, count * e_stride, hipMemcpyDeviceToHost, ctx->stream));
    CMD_HIP(hipStreamSynchronize(ctx->stream));
    return CMD_OK;
}

// Syndrome of n-bit words over the whole support (the decapsulation check's s_cmp): one key,
// tiles of 32 words in input order through the transposed-FFT check kernel.
extern "C" int cmd_stage_synd_full(cmd_ctx *ctx, int32_t key_id, size_t count, const uint8_t *r, size_t r_stride,
                                   uint16_t *s_out)
{
    if (!ctx || (count && (!r || !s_out))) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    int rc = check_key(ctx, key_id);
    if (rc) return rc;
    if (!count) return CMD_OK;
    if (count > (size_t)INT32_MAX) return CMD_ERR_ARG;
    CMD_HIP(hipSetDevice(ctx->device));
    const KeySlot &k = ctx->keys[key_id];
    const ParamDesc &p = kParams[k.pidx];
    if (r_stride < (size_t)p.n / 8) return CMD_ERR_ARG;
    std::vector<FusedTile> tiles;
    std::vector<int32_t> items(count), item_key(count, key_id);
    std::iota(items.begin(), items.end(), 0);
    for (size_t i = 0; i < count; i += 32) {
        FusedTile t;
        t.pidx = k.pidx;
        t.first = (int32_t)i;
        t.count = (int32_t)std::min<size_t>(32, count - i);
        t.pad = 0;
        tiles.push_back(t);
    }
    const size_t s_bytes = count * 2 * p.t * 2;
    const size_t off_key = (count * 4 + 255) / 256 * 256, off_tiles = off_key + off_key;
    rc = dev_reserve(ctx, ctx->e, count * r_stride);
    if (!rc) rc = dev_reserve(ctx, ctx->s, s_bytes);
    if (!rc) rc = dev_reserve(ctx, ctx->plan, off_tiles + tiles.size() * sizeof(FusedTile));
    if (rc) return rc;
    char *plan = (char *)ctx->plan.ptr;
    CMD_HIP(hipMemcpyAsync(ctx->e.ptr, r, count * r_stride, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(hipMemcpyAsync(plan, items.data(), count * 4, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(hipMemcpyAsync(plan + off_key, item_key.data(), count * 4, hipMemcpyHostToDevice, ctx->stream));
    CMD_HIP(hipMemcpyAsync(plan + off_tiles, tiles.data(), tiles.size() * sizeof(FusedTile), hipMemcpyHostToDevice,
                           ctx->stream));
    DecapArgs d;
    d.tiles = (const FusedTile *)(plan + off_tiles);
    d.n_tiles = (int32_t)tiles.size();
    d.pad0 = 0;
    d.items = (const int32_t *)plan;
    d.item_key = (const int32_t *)(plan + off_key);
    d.keys = ctx->d_keys;
    d.ct = nullptr;
    d.e = (const uint8_t *)ctx->e.ptr;
    d.ct_stride = 0;
    d.e_stride = r_stride;
    d.ok = nullptr;
    d.s_out = (uint16_t *)ctx->s.ptr;
    d.fft = ctx->d_fft;
    CMD_HIP(launch_decap_check(d, ctx->stream));
    CMD_HIP(hipMemcpyAsync(s_out, ctx->s.ptr, s_bytes, hipMemcpyDeviceToHost, ctx->stream));
    CMD_HIP(hipStreamSynchronize(ctx->stream));
    return CMD_OK;
}

// ---- schedules and timing --------------------------------------------------------------------

extern "C" int cmd_schedule_get(cmd_ctx *ctx, int32_t param_id, cmd_schedule *out)
{
    int pidx = param_index(param_id);
    if (!ctx || !out) return CMD_ERR_ARG;
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    const FusedSchedule &s = ctx->sched[pidx];
    out->tile_cts = s.tile_cts;
    out->waves_per_group = s.waves_per_group;
    out->waves_per_ct = s.waves_per_ct;
    out->groups_per_cu = s.groups_per_cu;
    out->root_alg = s.root_alg;
    return CMD_OK;
}

extern "C" int cmd_set_tile_fit(cmd_ctx *ctx, int32_t on)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    ctx->fit_tiles = on != 0;
    return CMD_OK;
}

extern "C" int cmd_schedule_set(cmd_ctx *ctx, int32_t param_id, const cmd_schedule *in)
{
    int pidx = param_index(param_id);
    if (!ctx || !in) return CMD_ERR_ARG;
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    FusedSchedule s{in->tile_cts, in->waves_per_group, in->waves_per_ct, in->groups_per_cu, in->root_alg};
    if (!fused_supported(s) || s.groups_per_cu < 1 || s.groups_per_cu > 16)
        return set_error(CMD_ERR_ARG, "schedule (%d, %d, %d, %d) not available", s.tile_cts, s.waves_per_group,
                         s.waves_per_ct, s.groups_per_cu);
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    ctx->sched[pidx] = s;
    return CMD_OK;
}

extern "C" int cmd_set_kernel_timing(cmd_ctx *ctx, int32_t on)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    ctx->timing = on != 0;
    ctx->timing_us = 0;
    ctx->timing_n = 0;
    return CMD_OK;
}

extern "C" int cmd_kernel_timing(cmd_ctx *ctx, double *avg_us, uint64_t *launches)
{
    if (!ctx) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    if (avg_us) *avg_us = ctx->timing_n ? ctx->timing_us / (double)ctx->timing_n : 0.0;
    if (launches) *launches = ctx->timing_n;
    ctx->timing_us = 0;
    ctx->timing_n = 0;
    return CMD_OK;
}

extern "C" const char *cmd_build_info(void)
{
    return "cm_decrypt__amd gfx950: plain(synd/bm/chien, uint16 GF) fused(tile 1..32, waves/group 1|2|4, one- or two-wave builds, "
           "bit-sliced BM + root search)";
}
