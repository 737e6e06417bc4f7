// Continuous-batching decryption server over the fused kernel.
//
// Requests (key id + ciphertext) queue on the host.  A worker thread keeps up to
// max_inflight launches on the device, each on its own stream with its own pinned staging
// and device buffers: whenever a launch slot is free and work is pending, the pending
// requests (any mix of parameter sets and resident keys) are planned into fused-kernel
// tiles and launched, so requests that arrive while others run on the GPU join the next
// launch and overlap with the running ones.  Finished launches are retired by polling
// their completion events; each request's error vector is copied out and its latency
// (submit -> result available) recorded.
//
// Host path: requests live in fixed blocks of a slot pool (never moved, reused through a
// free list); a request id carries its slot index and a generation number, so lookups need
// no map and a stale id is rejected.  The lock guards only the queue and the states: the
// worker copies ciphertexts into a launch and results out of it without holding it.
//
// Admission: keys are admitted by the context's HBM budget when loaded; a key with
// requests in flight cannot be unloaded; the in-flight I/O buffers are allocated up front
// within io_budget_bytes, and submit() rejects with CMD_ERR_FULL once queue_capacity
// requests are pending.  Every request's output lands in its own slot of its launch, so a
// result never depends on what else shares the launch.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>

#include "engine.h"

using namespace cmd;

namespace {

using Clock = std::chrono::steady_clock;

double now_us()
{
    return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

enum ReqState : uint8_t { kFree = 0, kPending = 1, kInFlight = 2, kDone = 3 };

struct Req {
    uint32_t gen = 0;
    uint8_t state = kFree;
    int32_t key = -1;
    int32_t pidx = -1;
    double t_submit = 0;
    uint8_t ct[kMaxSyndBytes];
    uint8_t e[kMaxEBytes];
};

constexpr int kBlockBits = 12;  // 4096 requests per pool block
constexpr uint32_t kBlockSize = 1u << kBlockBits;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint8_t *h_in = nullptr, *h_e = nullptr;  // pinned staging: [ciphertexts | work list], results
    DevBuf dev;                               // the same input image, then d_e
    std::vector<uint32_t> reqs;               // pool indices of the launch, in item order
    bool busy = false;                        // reserved by the worker (being launched or in flight)
    bool launched = false;                    // its completion event is recorded
};

}  // namespace

struct cmd_server {
    cmd_ctx *ctx = nullptr;
    cmd_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    // request pool: blocks are allocated on demand and never move
    std::vector<std::unique_ptr<Req[]>> blocks;
    std::vector<uint32_t> free_idx;
    uint32_t pool_size = 0, pool_cap = 0;
    std::deque<uint32_t> pending;
    std::vector<Slot> slots;
    std::thread worker;
    bool stop = false;
    int worker_rc = CMD_OK;
    // statistics
    uint64_t submitted = 0, completed = 0, rejected = 0, launches = 0, launched_items = 0;
    std::vector<double> lat_us;
    double busy_since = -1, busy_us = 0;
    int busy_slots = 0;
    size_t in_bytes = 0;  // capacity of a launch's input image

    Req &req(uint32_t idx) { return blocks[idx >> kBlockBits][idx & (kBlockSize - 1)]; }

    // A pool block with its pages already touched (allocated without mu held).
    static std::unique_ptr<Req[]> new_block()
    {
        std::unique_ptr<Req[]> blk(new (std::nothrow) Req[kBlockSize]);
        if (blk) {
            volatile char *p = reinterpret_cast<char *>(blk.get());
            for (size_t off = 0; off < sizeof(Req) * kBlockSize; off += 4096) p[off] = p[off];
        }
        return blk;
    }

    // Adds a block to the pool.  mu held.
    void install(std::unique_ptr<Req[]> blk)
    {
        blocks[pool_size >> kBlockBits] = std::move(blk);
        for (uint32_t i = kBlockSize; i-- > 0;) free_idx.push_back(pool_size + i);
        pool_size += kBlockSize;
    }

    // Pool index of a new request, or UINT32_MAX when the pool is exhausted.  mu held.
    uint32_t alloc()
    {
        if (free_idx.empty()) {
            if (pool_size >= pool_cap) return UINT32_MAX;
            std::unique_ptr<Req[]> blk(new (std::nothrow) Req[kBlockSize]);
            if (!blk) return UINT32_MAX;
            install(std::move(blk));
        }
        const uint32_t idx = free_idx.back();
        free_idx.pop_back();
        return idx;
    }

    // The request an id names, or nullptr for an unknown or already taken id.  mu held.
    Req *lookup(uint64_t id)
    {
        const uint32_t idx = (uint32_t)id, gen = (uint32_t)(id >> 32);
        if (idx >= pool_size) return nullptr;
        Req &r = req(idx);
        return (r.state != kFree && r.gen == gen) ? &r : nullptr;
    }
};

namespace {

constexpr size_t kCtStride = kMaxSyndBytes;
constexpr size_t kEStride = kMaxEBytes;

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Plans and enqueues the launch of slot s (its reqs already chosen).  Runs without srv->mu:
// in-flight requests are not touched by other threads.
int launch_slot(cmd_server *srv, Slot &s)
{
    cmd_ctx *ctx = srv->ctx;
    const size_t count = s.reqs.size();
    std::vector<int32_t> key_ids(count);
    for (size_t i = 0; i < count; i++) key_ids[i] = srv->req(s.reqs[i]).key;
    std::vector<FusedTile> tiles;
    std::vector<int32_t> items, item_key;
    int launch_tile = 32;
    FusedSchedule sched;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->lock);
        // Latency first: narrow the tiles until the launch has about one wavefront per SIMD
        // (a tile's Berlekamp-Massey takes the same time for 1 or 32 ciphertexts, while its
        // syndrome and root search grow with the count), never wider than the tuned width.
        FusedSchedule per_set[kNumParams];
        const int target = ctx->n_cu * 4;
        for (int i = 0; i < kNumParams; i++) {
            per_set[i] = ctx->sched[i];
            int tile = per_set[i].tile_cts;
            while (tile > 1 && (int)((count + tile - 1) / tile) < target) tile >>= 1;
            per_set[i].tile_cts = tile;
        }
        plan_fused(ctx, per_set, count, key_ids.data(), tiles, items, item_key, &launch_tile);
        sched = launch_schedule(per_set, tiles, launch_tile);
    }
    // one input image per launch, copied in one transfer: [ct rows | items | item keys |
    // tiles | work-queue counter (zero)]
    const size_t off_items = align256(count * kCtStride);
    const size_t off_keys = off_items + align256(count * 4), off_tiles = off_keys + align256(count * 4);
    const size_t off_ctr = off_tiles + align256(tiles.size() * sizeof(FusedTile));
    const size_t in_bytes = off_ctr + 4;
    for (size_t i = 0; i < count; i++) {
        const Req &r = srv->req(s.reqs[i]);
        std::memcpy(s.h_in + i * kCtStride, r.ct, kParams[r.pidx].synd_bytes);
    }
    std::memcpy(s.h_in + off_items, items.data(), items.size() * 4);
    std::memcpy(s.h_in + off_keys, item_key.data(), item_key.size() * 4);
    std::memcpy(s.h_in + off_tiles, tiles.data(), tiles.size() * sizeof(FusedTile));
    std::memset(s.h_in + off_ctr, 0, 4);
    char *d_in = (char *)s.dev.ptr;
    uint8_t *d_e = (uint8_t *)(d_in + srv->in_bytes);
    CMD_HIP(hipMemcpyAsync(d_in, s.h_in, in_bytes, hipMemcpyHostToDevice, s.stream));
    FusedArgs a;
    a.tiles = (const FusedTile *)(d_in + off_tiles);
    a.n_tiles = (int32_t)tiles.size();
    a.pad0 = 0;
    a.items = (const int32_t *)(d_in + off_items);
    a.item_key = (const int32_t *)(d_in + off_keys);
    a.keys = ctx->d_keys;
    a.ct = (const uint8_t *)d_in;
    a.e = d_e;
    a.ct_stride = kCtStride;
    a.e_stride = kEStride;
    a.tile_counter = (uint32_t *)(d_in + off_ctr);
    a.fft = ctx->d_fft;
    CMD_HIP(launch_fused(a, sched, ctx->n_cu, s.stream));
    CMD_HIP(hipMemcpyAsync(s.h_e, d_e, count * kEStride, hipMemcpyDeviceToHost, s.stream));
    CMD_HIP(hipEventRecord(s.done, s.stream));
    return CMD_OK;
}

// Copies a finished launch's results into its requests (without srv->mu), then marks them
// done and records statistics (with it).
void retire_slot(cmd_server *srv, Slot &s, std::unique_lock<std::mutex> &lk, bool ok)
{
    lk.unlock();
    const double t = now_us();
    if (ok)
        for (size_t i = 0; i < s.reqs.size(); i++) {
            Req &r = srv->req(s.reqs[i]);
            std::memcpy(r.e, s.h_e + i * kEStride, kParams[r.pidx].n / 8);
        }
    {
        std::lock_guard<std::recursive_mutex> g(srv->ctx->lock);
        for (uint32_t idx : s.reqs) srv->ctx->keys[srv->req(idx).key].inflight--;
    }
    lk.lock();
    for (uint32_t idx : s.reqs) {
        Req &r = srv->req(idx);
        r.state = kDone;
        srv->lat_us.push_back(t - r.t_submit);
        srv->completed++;
    }
    s.reqs.clear();
    s.busy = false;
    s.launched = false;
    if (--srv->busy_slots == 0) {
        srv->busy_us += t - srv->busy_since;
        srv->busy_since = -1;
    }
}

// Marks queued requests finished without a result (after a device error); waiters then get
// worker_rc.  mu held.
void fail_pending(cmd_server *srv)
{
    std::lock_guard<std::recursive_mutex> g(srv->ctx->lock);
    for (uint32_t idx : srv->pending) {
        Req &r = srv->req(idx);
        srv->ctx->keys[r.key].inflight--;
        r.state = kDone;
    }
    srv->pending.clear();
}

void worker_main(cmd_server *srv)
{
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        bool progressed = false;
        for (Slot &s : srv->slots) {
            if (!s.busy || !s.launched) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipSuccess) {
                retire_slot(srv, s, lk, true);
                progressed = true;
            } else if (q != hipErrorNotReady) {
                srv->worker_rc = hip_fail(q, "server launch");
                retire_slot(srv, s, lk, false);  // results are invalid; waiters see the error code
                progressed = true;
            }
        }
        if (progressed) srv->cv_done.notify_all();

        if (!srv->pending.empty() && srv->worker_rc == CMD_OK) {
            Slot *free_slot = nullptr;
            for (Slot &s : srv->slots)
                if (!s.busy) { free_slot = &s; break; }
            const double waited = now_us() - srv->req(srv->pending.front()).t_submit;
            const bool ready = (int)srv->pending.size() >= srv->cfg.min_batch || waited >= srv->cfg.max_wait_us ||
                               srv->busy_slots == 0 || srv->stop;
            if (free_slot && ready) {
                Slot &s = *free_slot;
                while (!srv->pending.empty() && (int)s.reqs.size() < srv->cfg.max_batch) {
                    const uint32_t idx = srv->pending.front();
                    srv->pending.pop_front();
                    srv->req(idx).state = kInFlight;
                    s.reqs.push_back(idx);
                }
                s.busy = true;
                const double t = now_us();
                if (srv->busy_slots++ == 0) srv->busy_since = t;
                srv->launches++;
                srv->launched_items += s.reqs.size();
                lk.unlock();
                const int rc = launch_slot(srv, s);
                lk.lock();
                if (rc) {
                    srv->worker_rc = rc;
                    retire_slot(srv, s, lk, false);
                    srv->cv_done.notify_all();
                } else {
                    s.launched = true;
                }
                progressed = true;
            }
        }
        if (srv->worker_rc != CMD_OK && !srv->pending.empty()) {
            fail_pending(srv);
            srv->cv_done.notify_all();
        }
        if (srv->stop && srv->pending.empty() && srv->busy_slots == 0) break;
        if (!progressed) {
            // poll launches in flight often; otherwise sleep until a submit arrives
            if (srv->busy_slots > 0 || !srv->pending.empty())
                srv->cv_work.wait_for(lk, std::chrono::microseconds(10));
            else
                srv->cv_work.wait(lk);
        }
    }
}

}  // namespace

extern "C" int cmd_server_create(cmd_ctx *ctx, const cmd_server_config *cfg, cmd_server **out)
{
    if (!ctx || !cfg || !out) return CMD_ERR_ARG;
    *out = nullptr;
    cmd_server_config c = *cfg;
    if (c.max_batch < 1) c.max_batch = 4096;
    if (c.max_inflight < 1) c.max_inflight = 2;
    if (c.max_inflight > 8) c.max_inflight = 8;
    if (c.queue_capacity < 1) c.queue_capacity = 1 << 20;
    if (c.min_batch < 1) c.min_batch = 1;
    if (c.max_wait_us < 0) c.max_wait_us = 0;
    const size_t mb = (size_t)c.max_batch;
    const size_t in_bytes = align256(mb * kCtStride) + align256(mb * 4) * 2 + align256(mb * sizeof(FusedTile)) + 256;
    const size_t per_slot = in_bytes + align256(mb * kEStride);
    if (c.io_budget_bytes && per_slot * (size_t)c.max_inflight > c.io_budget_bytes)
        return set_error(CMD_ERR_BUDGET, "%d launch slots of %zu bytes exceed the I/O budget of %llu bytes",
                         c.max_inflight, per_slot, (unsigned long long)c.io_budget_bytes);

    std::unique_ptr<cmd_server> srv(new cmd_server());
    srv->ctx = ctx;
    srv->cfg = c;
    srv->in_bytes = in_bytes;
    // pending requests plus those of every launch slot, rounded up to whole blocks
    const uint64_t cap = (uint64_t)c.queue_capacity + (uint64_t)c.max_inflight * (uint64_t)c.max_batch;
    const uint64_t nblocks = (cap + kBlockSize - 1) / kBlockSize;
    if (nblocks > (UINT32_MAX >> kBlockBits)) return set_error(CMD_ERR_ARG, "queue capacity too large");
    srv->blocks.resize(nblocks);
    srv->pool_cap = (uint32_t)(nblocks << kBlockBits);
    srv->slots.resize(c.max_inflight);
    for (Slot &s : srv->slots) s.reqs.reserve(mb);
    srv->lat_us.reserve(1 << 18);
    // a first few blocks up front; submit grows the pool ahead of need
    for (int b = 0; b < 4 && srv->pool_size < srv->pool_cap; b++) {
        std::unique_ptr<Req[]> blk = cmd_server::new_block();
        if (!blk) break;
        srv->install(std::move(blk));
    }
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    int rc = CMD_OK;
    for (Slot &s : srv->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&s.h_in, in_bytes, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s.h_e, mb * kEStride, hipHostMallocDefault) != hipSuccess) {
            rc = set_error(CMD_ERR_HIP, "server slot allocation failed");
            break;
        }
        rc = dev_alloc(ctx, s.dev, per_slot);
        if (rc) break;
    }
    if (rc) {
        for (Slot &s : srv->slots) {
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.h_in) (void)hipHostFree(s.h_in);
            if (s.h_e) (void)hipHostFree(s.h_e);
            dev_free(ctx, s.dev);
        }
        return rc;
    }
    cmd_server *raw = srv.get();
    srv->worker = std::thread([raw]() { worker_main(raw); });
    *out = srv.release();
    return CMD_OK;
}

extern "C" int cmd_server_submit_batch(cmd_server *s, size_t count, const int32_t *key_ids, const uint8_t *ct,
                                       size_t ct_stride, uint64_t *reqs, size_t *accepted)
{
    if (accepted) *accepted = 0;
    if (!s || (count && (!key_ids || !ct || !reqs))) return CMD_ERR_ARG;
    std::vector<int32_t> pidx(count);
    {
        std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
        for (size_t i = 0; i < count; i++) {
            const int32_t k = key_ids[i];
            if (k < 0 || (size_t)k >= s->ctx->keys.size() || !s->ctx->keys[k].used)
                return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)k);
            pidx[i] = s->ctx->keys[k].pidx;
            if (ct_stride < (size_t)kParams[pidx[i]].synd_bytes)
                return set_error(CMD_ERR_ARG, "ct_stride %zu below the ciphertext size of item %zu", ct_stride, i);
        }
    }
    // Grow the pool ahead of need without holding mu: a block is 4096 requests (5 MB), and
    // allocating it under the lock would stall the worker's launches and retirements.
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->free_idx.size() >= count + kBlockSize / 2 || s->pool_size >= s->pool_cap) break;
        }
        std::unique_ptr<Req[]> blk = cmd_server::new_block();
        if (!blk) break;
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->pool_size < s->pool_cap) s->install(std::move(blk));
    }
    size_t done = 0;
    int rc = CMD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return set_error(CMD_ERR_ARG, "server is shutting down");
        const double t = now_us();
        for (; done < count; done++) {
            if ((int)s->pending.size() >= s->cfg.queue_capacity) {
                s->rejected += count - done;
                rc = set_error(CMD_ERR_FULL, "%d requests pending", (int)s->pending.size());
                break;
            }
            const uint32_t idx = s->alloc();
            if (idx == UINT32_MAX) {
                rc = set_error(CMD_ERR_NOMEM, "request pool exhausted");
                break;
            }
            Req &r = s->req(idx);
            r.gen++;
            r.state = kPending;
            r.key = key_ids[done];
            r.pidx = pidx[done];
            r.t_submit = t;
            std::memcpy(r.ct, ct + done * ct_stride, kParams[r.pidx].synd_bytes);
            s->pending.push_back(idx);
            s->submitted++;
            reqs[done] = ((uint64_t)r.gen << 32) | idx;
        }
        if (done) {
            // held against cmd_key_unload until retired (context lock after the server lock,
            // the order every path here uses)
            std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
            for (size_t i = 0; i < done; i++) s->ctx->keys[key_ids[i]].inflight++;
        }
    }
    if (accepted) *accepted = done;
    if (done) s->cv_work.notify_one();
    return rc;
}

extern "C" int cmd_server_submit(cmd_server *s, int32_t key_id, const uint8_t *ct, uint64_t *req)
{
    if (!s || !ct || !req) return CMD_ERR_ARG;
    return cmd_server_submit_batch(s, 1, &key_id, ct, kMaxSyndBytes, req, nullptr);
}

static int take_result(cmd_server *s, uint64_t id, uint8_t *e_out)
{
    Req *r = s->lookup(id);
    if (!r) return set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    if (r->state != kDone) return CMD_ERR_PENDING;
    if (s->worker_rc) return s->worker_rc;
    if (e_out) std::memcpy(e_out, r->e, kParams[r->pidx].n / 8);
    r->state = kFree;
    s->free_idx.push_back((uint32_t)id);
    return CMD_OK;
}

extern "C" int cmd_server_poll(cmd_server *s, uint64_t req, uint8_t *e_out)
{
    if (!s) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    return take_result(s, req, e_out);
}

extern "C" int cmd_server_wait(cmd_server *s, uint64_t req, uint8_t *e_out, double timeout_us)
{
    if (!s) return CMD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = Clock::now() + std::chrono::microseconds((long long)(timeout_us < 0 ? 0 : timeout_us));
    for (;;) {
        int rc = take_result(s, req, e_out);
        if (rc != CMD_ERR_PENDING) return rc;
        if (timeout_us < 0) {
            s->cv_done.wait(lk);
        } else if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            rc = take_result(s, req, e_out);
            return rc == CMD_ERR_PENDING ? CMD_ERR_TIMEOUT : rc;
        }
    }
}

extern "C" int cmd_server_wait_batch(cmd_server *s, size_t count, const uint64_t *reqs, uint8_t *e_out,
                                     size_t e_stride, double timeout_us)
{
    if (!s || (count && (!reqs || !e_out))) return CMD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = Clock::now() + std::chrono::microseconds((long long)(timeout_us < 0 ? 0 : timeout_us));
    for (size_t i = 0; i < count; i++) {
        const Req *r = s->lookup(reqs[i]);
        if (!r) return set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)reqs[i]);
        if ((size_t)kParams[r->pidx].n / 8 > e_stride)
            return set_error(CMD_ERR_ARG, "e_stride %zu below the error vector of request %zu", e_stride, i);
    }
    for (size_t i = 0; i < count; i++) {
        for (;;) {
            int rc = take_result(s, reqs[i], e_out + i * e_stride);
            if (rc != CMD_ERR_PENDING) {
                if (rc) return rc;
                break;
            }
            if (timeout_us < 0) {
                s->cv_done.wait(lk);
            } else if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
                rc = take_result(s, reqs[i], e_out + i * e_stride);
                if (rc == CMD_ERR_PENDING) return CMD_ERR_TIMEOUT;
                if (rc) return rc;
                break;
            }
        }
    }
    return CMD_OK;
}

extern "C" int cmd_server_stats_get(cmd_server *s, cmd_server_stats *out)
{
    if (!s || !out) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->rejected = s->rejected;
    out->launches = s->launches;
    out->mean_batch = s->launches ? (double)s->launched_items / (double)s->launches : 0.0;
    std::vector<double> v = s->lat_us;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        size_t i = (size_t)(q * (double)(v.size() - 1) + 0.5);
        return v[std::min(i, v.size() - 1)];
    };
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    out->max_us = v.empty() ? 0.0 : v.back();
    double busy = s->busy_us;
    if (s->busy_since >= 0) busy += now_us() - s->busy_since;
    out->busy_s = busy * 1e-6;
    return CMD_OK;
}

extern "C" int cmd_server_reset_stats(cmd_server *s)
{
    if (!s) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->rejected = s->launches = s->launched_items = 0;
    s->lat_us.clear();
    s->busy_us = 0;
    if (s->busy_since >= 0) s->busy_since = now_us();
    return CMD_OK;
}

extern "C" int cmd_server_destroy(cmd_server *s)
{
    if (!s) return CMD_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
    (void)hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        if (sl.h_e) (void)hipHostFree(sl.h_e);
        dev_free(s->ctx, sl.dev);
    }
    delete s;  // every request was retired by the worker; the pool blocks go with the server
    return CMD_OK;
}
