// Continuous-batching server over the fused kernel: decryptions and decapsulations.
//
// Requests (key id + ciphertext, and whether the caller wants the error vector or the KEM
// session key) queue on the host.  A worker thread keeps up to max_inflight launches on the
// device, each on its own stream with its own pinned staging and device buffers: whenever a
// launch slot is free and work is pending, the pending requests (any mix of parameter sets,
// resident keys and request kinds) are planned into fused-kernel tiles and launched, so
// requests that arrive while others run on the GPU join the next launch and overlap with the
// running ones.  A launch's decapsulations sit in tiles of their own after its decryptions;
// the re-encryption check and SHAKE256 run over those tiles after the fused kernel.
// Finished launches are retired by polling their completion events; each request's result
// is copied out and its latency (submit -> result available) recorded, split into the wait
// for a launch and the launch itself.
//
// Host path: requests live in fixed blocks of a slot pool (never moved, reused through a
// free list); a request id carries its slot index and a generation number, so lookups need
// no map and a stale id is rejected.  The lock guards only the queue and the states: the
// worker copies ciphertexts into a launch and results out of it without holding it.
//
// Admission: keys are admitted by the context's HBM budget when loaded.  Submission checks
// a request's key and holds it against cmd_key_unload in one section under the context
// lock, so a key cannot be unloaded between the check and the launch; the hold ends when the
// request retires.  The in-flight I/O buffers are allocated up front within
// io_budget_bytes, and submit() rejects with CMD_ERR_FULL once queue_capacity requests are
// pending.  Every request's output lands in its own row of its launch, so a result never
// depends on what else shares the launch.  A launch that fails (a HIP error, or the device
// scratch of its decapsulations not fitting the budget) fails its own requests only.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>

#include <sys/prctl.h>

#include "engine.h"
#include "trace.h"

using namespace cmd;

namespace {

using Clock = std::chrono::steady_clock;

double now_us()
{
    return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

// kTaken: finished and being copied out by cmd_server_wait_batch without the lock held
enum ReqState : uint8_t { kFree = 0, kPending = 1, kInFlight = 2, kDone = 3, kTaken = 4 };

struct Req {
    uint32_t gen = 0;
    uint8_t state = kFree;
    uint8_t mode = CMD_MODE_FUSED;
    uint16_t out_bytes = 0;  // result size: n/8 (decryption) or 32 (decapsulation)
    int32_t key = -1;
    int32_t pidx = -1;
    int32_t rc = CMD_OK;     // CMD_OK, or the error of the launch that carried the request
    double t_submit = 0, t_launch = 0;
    uint8_t ct[kMaxSyndBytes];
    uint8_t e[kMaxEBytes];
};

constexpr int kBlockBits = 12;  // 4096 requests per pool block
constexpr uint32_t kBlockSize = 1u << kBlockBits;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, done = nullptr;  // around the launch's device work (timing events)
    uint8_t *h_in = nullptr, *h_e = nullptr;  // pinned staging: [ciphertexts | work list], results
    uint8_t *h_e_dev = nullptr;               // h_e as the device addresses it
    DevBuf dev;                               // the same input image, then d_e
    DevBuf kem;                               // decapsulation scratch (accept flags, SHAKE256 inputs)
    std::vector<uint32_t> reqs;               // pool indices of the launch, in item order
    size_t n_launch = 0;                      // leading reqs that are in the launch (the rest failed)
    bool busy = false;                        // reserved by the worker (being launched or in flight)
    bool launched = false;                    // its completion event is recorded
    double t_launch = 0, t_enq = 0, t_seen = 0;  // popped, enqueued, completion seen (host clock)
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
    // statistics
    uint64_t submitted = 0, completed = 0, rejected = 0, launches = 0, launched_items = 0, failed = 0;
    std::vector<double> lat_us, queue_us, dev_us, gpu_us;
    std::vector<cmd_server_launch> launch_log;
    double t_created = 0;
    double step_max[4] = {0, 0, 0, 0};  // launch, retire, poll, wait overshoot (us)
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
// in-flight requests are not touched by other threads.  Requests whose key is no longer the
// one they were submitted for (which the hold taken at submission rules out) are moved to
// the end of s.reqs with CMD_ERR_KEY and left out of the launch.
int launch_slot(cmd_server *srv, Slot &s)
{
    TraceRange tr("server launch");
    cmd_ctx *ctx = srv->ctx;
    std::vector<uint32_t> dec, kem, bad;
    dec.reserve(s.reqs.size());
    std::vector<FusedTile> tiles, tiles_k;
    std::vector<int32_t> items, item_key, items_k, item_key_k;
    int launch_tile = 32, launch_tile_k = 0;
    FusedSchedule sched;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->lock);
        for (uint32_t idx : s.reqs) {
            Req &r = srv->req(idx);
            const bool key_ok = r.key >= 0 && (size_t)r.key < ctx->keys.size() && ctx->keys[r.key].used &&
                                ctx->keys[r.key].pidx == r.pidx;
            if (!key_ok) {
                r.rc = set_error(CMD_ERR_KEY, "key id %d no longer holds the request's key", (int)r.key);
                bad.push_back(idx);
            } else {
                (r.mode == CMD_MODE_KEM ? kem : dec).push_back(idx);
            }
        }
        s.reqs.clear();
        s.reqs.insert(s.reqs.end(), dec.begin(), dec.end());
        s.reqs.insert(s.reqs.end(), kem.begin(), kem.end());
        s.n_launch = s.reqs.size();
        s.reqs.insert(s.reqs.end(), bad.begin(), bad.end());
        if (s.n_launch == 0) return CMD_OK;
        // Latency first: narrow the tiles until the launch has about one wavefront per SIMD
        // (a tile's Berlekamp-Massey takes the same time for 1 or 32 ciphertexts, while its
        // syndrome and root search grow with the count), never wider than the tuned width.
        const size_t count = s.n_launch;
        FusedSchedule per_set[kNumParams];
        const int target = ctx->n_cu * 4;
        for (int i = 0; i < kNumParams; i++) {
            per_set[i] = ctx->sched[i];
            int tile = per_set[i].tile_cts;
            while (tile > 1 && (int)((count + tile - 1) / tile) < target) tile >>= 1;
            per_set[i].tile_cts = tile;
        }
        // decryptions first, then decapsulations: tiles of one kind each, so the check and the
        // session keys run over a contiguous range of tiles and items
        std::vector<int32_t> keys_d(dec.size()), keys_k(kem.size());
        for (size_t i = 0; i < dec.size(); i++) keys_d[i] = srv->req(dec[i]).key;
        for (size_t i = 0; i < kem.size(); i++) keys_k[i] = srv->req(kem[i]).key;
        int rc = plan_fused(ctx, per_set, dec.size(), keys_d.data(), tiles, items, item_key, &launch_tile);
        if (!rc && !kem.empty())
            rc = plan_fused(ctx, per_set, kem.size(), keys_k.data(), tiles_k, items_k, item_key_k, &launch_tile_k);
        if (rc) return rc;
        if (dec.empty()) launch_tile = launch_tile_k;
        launch_tile = std::max(launch_tile, launch_tile_k);
        const int32_t item_base = (int32_t)items.size();
        for (FusedTile t : tiles_k) {
            t.first += item_base;
            tiles.push_back(t);
        }
        for (size_t i = 0; i < items_k.size(); i++) {
            items.push_back(items_k[i] + item_base);
            item_key.push_back(item_key_k[i]);
        }
        sched = launch_schedule(per_set, tiles, launch_tile);
        if (!kem.empty()) {
            rc = dev_reserve(ctx, s.kem, align256(count * 4) + kem.size() * (size_t)kKemMsgStride);
            if (rc) return rc;
        }
    }
    const size_t count = s.n_launch;
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
    // A launch of decryptions only writes its error vectors straight into the pinned result
    // buffer (no copy back after the kernel); decapsulations read e again on the device, so a
    // launch with any keeps e in device memory and copies the results back.
    const bool direct = kem.empty();
    CMD_HIP(hipEventRecord(s.start, s.stream));
    CMD_HIP(hipMemcpyAsync(d_in, s.h_in, in_bytes, hipMemcpyHostToDevice, s.stream));
    FusedArgs a;
    a.tiles = (const FusedTile *)(d_in + off_tiles);
    a.n_tiles = (int32_t)tiles.size();
    a.pad0 = 0;
    a.items = (const int32_t *)(d_in + off_items);
    a.item_key = (const int32_t *)(d_in + off_keys);
    a.keys = ctx->d_keys;
    a.ct = (const uint8_t *)d_in;
    a.e = direct ? s.h_e_dev : d_e;
    a.ct_stride = kCtStride;
    a.e_stride = kEStride;
    a.tile_counter = (uint32_t *)(d_in + off_ctr);
    a.fft = ctx->d_fft;
    CMD_HIP(launch_fused(a, sched, ctx->n_cu, s.stream));
    if (!kem.empty()) {
        // the check reads each decapsulation's decoded e from its row; the session key then
        // replaces the row's first 32 bytes (k_kem_pack has copied e out before k_keccak runs)
        const int32_t n_tiles_d = (int32_t)(tiles.size() - tiles_k.size());
        const int32_t n_items_d = (int32_t)dec.size();
        uint32_t *d_ok = (uint32_t *)s.kem.ptr;
        uint8_t *d_msg = (uint8_t *)s.kem.ptr + align256(count * 4);
        DecapArgs d;
        d.tiles = a.tiles + n_tiles_d;
        d.n_tiles = (int32_t)tiles_k.size();
        d.pad0 = 0;
        d.items = a.items;
        d.item_key = a.item_key;
        d.keys = ctx->d_keys;
        d.ct = a.ct;
        d.e = d_e;
        d.ct_stride = kCtStride;
        d.e_stride = kEStride;
        d.ok = d_ok;
        d.s_out = nullptr;
        d.fft = ctx->d_fft;
        CMD_HIP(launch_decap_check(d, s.stream));
        KemHashArgs h;
        h.count = (int32_t)kem.size();
        h.pad0 = 0;
        h.items = a.items + n_items_d;
        h.item_key = a.item_key + n_items_d;
        h.keys = ctx->d_keys;
        h.ct = a.ct;
        h.e = d_e;
        h.ct_stride = kCtStride;
        h.e_stride = kEStride;
        h.ok = d_ok;
        h.msg = d_msg;
        h.key_out = d_e;
        h.key_stride = kEStride;
        CMD_HIP(launch_kem_hash(h, s.stream));
    }
    if (!direct) CMD_HIP(hipMemcpyAsync(s.h_e, d_e, count * kEStride, hipMemcpyDeviceToHost, s.stream));
    CMD_HIP(hipEventRecord(s.done, s.stream));
    return CMD_OK;
}

// Copies a finished launch's results into its requests (without srv->mu), then marks them
// done and records statistics (with it).  launch_rc != CMD_OK fails every request of the
// launch; requests already marked with an error keep theirs.
void retire_slot(cmd_server *srv, Slot &s, std::unique_lock<std::mutex> &lk, int launch_rc)
{
    lk.unlock();
    const double t = now_us();
    float gpu_ms = -1.f;
    if (launch_rc == CMD_OK && s.n_launch && hipEventElapsedTime(&gpu_ms, s.start, s.done) != hipSuccess)
        gpu_ms = -1.f;
    for (size_t i = 0; i < s.reqs.size(); i++) {
        Req &r = srv->req(s.reqs[i]);
        if (r.rc == CMD_OK && launch_rc != CMD_OK) r.rc = launch_rc;
        if (r.rc == CMD_OK && i < s.n_launch) std::memcpy(r.e, s.h_e + i * kEStride, r.out_bytes);
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
        srv->queue_us.push_back(r.t_launch - r.t_submit);
        srv->completed++;
        if (r.rc != CMD_OK) srv->failed++;
    }
    if (launch_rc == CMD_OK && s.n_launch) srv->dev_us.push_back(t - s.t_launch);
    if (gpu_ms >= 0.f) {
        srv->gpu_us.push_back(1e3 * gpu_ms);
        cmd_server_launch rec;
        rec.t_us = s.t_launch - srv->t_created;
        rec.enqueue_us = s.t_enq - s.t_launch;
        rec.gpu_us = 1e3 * gpu_ms;
        rec.seen_us = s.t_seen - s.t_enq;
        rec.retire_us = now_us() - s.t_seen;
        rec.items = (uint32_t)s.n_launch;
        rec.pad = 0;
        srv->launch_log.push_back(rec);
    }
    s.reqs.clear();
    s.n_launch = 0;
    s.busy = false;
    s.launched = false;
    if (--srv->busy_slots == 0) {
        srv->busy_us += t - srv->busy_since;
        srv->busy_since = -1;
    }
}

// The longest worker step of each kind since the last reset.  mu held.
void note_step(cmd_server *srv, int kind, double us)
{
    if (us > srv->step_max[kind]) srv->step_max[kind] = us;
}

constexpr int kPollUs = 10;  // completion polling interval while launches are in flight

void worker_main(cmd_server *srv)
{
    (void)hipSetDevice(srv->ctx->device);
    // the completion polls sleep kPollUs; the default timer slack (50 us) would stretch each
    prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        bool progressed = false;
        for (Slot &s : srv->slots) {
            if (!s.busy || !s.launched) continue;
            const double tq = now_us();
            hipError_t q = hipEventQuery(s.done);
            const double tr = now_us();
            note_step(srv, 2, tr - tq);
            if (q == hipSuccess) {
                s.t_seen = tr;
                retire_slot(srv, s, lk, CMD_OK);
                note_step(srv, 1, now_us() - tr);
                progressed = true;
            } else if (q != hipErrorNotReady) {
                const int rc = hip_fail(q, "server launch");
                lk.unlock();
                (void)hipStreamSynchronize(s.stream);  // nothing of the launch may still touch its buffers
                lk.lock();
                retire_slot(srv, s, lk, rc);
                progressed = true;
            }
        }
        if (progressed) srv->cv_done.notify_all();

        if (!srv->pending.empty()) {
            Slot *free_slot = nullptr;
            for (Slot &s : srv->slots)
                if (!s.busy) { free_slot = &s; break; }
            const double waited = now_us() - srv->req(srv->pending.front()).t_submit;
            const bool ready = (int)srv->pending.size() >= srv->cfg.min_batch || waited >= srv->cfg.max_wait_us ||
                               srv->busy_slots == 0 || srv->stop;
            if (free_slot && ready) {
                Slot &s = *free_slot;
                const double t = now_us();
                while (!srv->pending.empty() && (int)s.reqs.size() < srv->cfg.max_batch) {
                    const uint32_t idx = srv->pending.front();
                    srv->pending.pop_front();
                    Req &r = srv->req(idx);
                    r.state = kInFlight;
                    r.t_launch = t;
                    s.reqs.push_back(idx);
                }
                s.busy = true;
                s.t_launch = t;
                if (srv->busy_slots++ == 0) srv->busy_since = t;
                srv->launches++;
                srv->launched_items += s.reqs.size();
                lk.unlock();
                int rc = launch_slot(srv, s);
                s.t_enq = now_us();
                if (rc) (void)hipStreamSynchronize(s.stream);  // copies already queued finish first
                lk.lock();
                note_step(srv, 0, now_us() - t);
                if (rc || s.n_launch == 0) {
                    retire_slot(srv, s, lk, rc);
                    srv->cv_done.notify_all();
                } else {
                    s.launched = true;
                }
                progressed = true;
            }
        }
        if (srv->stop && srv->pending.empty() && srv->busy_slots == 0) break;
        if (!progressed) {
            // poll launches in flight often; otherwise sleep until a submit arrives
            if (srv->busy_slots > 0 || !srv->pending.empty()) {
                const double tw = now_us();
                srv->cv_work.wait_for(lk, std::chrono::microseconds(kPollUs));
                note_step(srv, 3, now_us() - tw - kPollUs);
            } else
                srv->cv_work.wait(lk);
        }
    }
}

int submit_impl(cmd_server *s, int32_t mode, size_t count, const int32_t *key_ids, const uint8_t *ct,
                size_t ct_stride, uint64_t *reqs, size_t *accepted)
{
    if (accepted) *accepted = 0;
    if (!s || (count && (!key_ids || !ct || !reqs))) return CMD_ERR_ARG;
    if (mode != CMD_MODE_FUSED && mode != CMD_MODE_KEM)
        return set_error(CMD_ERR_ARG, "server requests are CMD_MODE_FUSED or CMD_MODE_KEM, not %d", (int)mode);
    if (!count) return CMD_OK;
    std::vector<int32_t> pidx(count);
    {
        // check every key and hold it in one section: cmd_key_unload cannot slip in between
        std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
        for (size_t i = 0; i < count; i++) {
            const int32_t k = key_ids[i];
            if (k < 0 || (size_t)k >= s->ctx->keys.size() || !s->ctx->keys[k].used)
                return set_error(CMD_ERR_KEY, "item %zu: key id %d is not loaded", i, (int)k);
            pidx[i] = s->ctx->keys[k].pidx;
            if (ct_stride < (size_t)kParams[pidx[i]].synd_bytes)
                return set_error(CMD_ERR_ARG, "ct_stride %zu below the ciphertext size of item %zu", ct_stride, i);
        }
        for (size_t i = 0; i < count; i++) s->ctx->keys[key_ids[i]].inflight++;
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
        if (s->stop) {
            rc = set_error(CMD_ERR_ARG, "server is shutting down");
        } else {
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
                r.mode = (uint8_t)mode;
                r.key = key_ids[done];
                r.pidx = pidx[done];
                r.out_bytes = (uint16_t)(mode == CMD_MODE_KEM ? CMD_KEM_KEY_BYTES : kParams[r.pidx].n / 8);
                r.rc = CMD_OK;
                r.t_submit = t;
                r.t_launch = t;
                std::memcpy(r.ct, ct + done * ct_stride, kParams[r.pidx].synd_bytes);
                s->pending.push_back(idx);
                s->submitted++;
                reqs[done] = ((uint64_t)r.gen << 32) | idx;
            }
        }
    }
    if (done < count) {
        // release the holds of the requests that were not queued
        std::lock_guard<std::recursive_mutex> g(s->ctx->lock);
        for (size_t i = done; i < count; i++) s->ctx->keys[key_ids[i]].inflight--;
    }
    if (accepted) *accepted = done;
    if (done) s->cv_work.notify_one();
    return rc;
}

// Hands out a finished request's result and frees it.  mu held.
int take_result(cmd_server *s, uint64_t id, uint8_t *out, size_t cap)
{
    Req *r = s->lookup(id);
    if (!r || r->state == kTaken) return set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)id);
    if (r->state != kDone) return CMD_ERR_PENDING;
    const int rc = r->rc;
    if (rc == CMD_OK && out) {
        if (cap < r->out_bytes)
            return set_error(CMD_ERR_ARG, "result buffer of %zu bytes, %d needed", cap, (int)r->out_bytes);
        std::memcpy(out, r->e, r->out_bytes);
    }
    r->state = kFree;
    s->free_idx.push_back((uint32_t)id);
    return rc;
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t i = (size_t)(q * (double)(v.size() - 1) + 0.5);
    return v[std::min(i, v.size() - 1)];
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
    // sample vectors sized for long measurement windows: growing them under mu would stall
    // the worker
    srv->lat_us.reserve(1 << 20);
    srv->queue_us.reserve(1 << 20);
    srv->dev_us.reserve(1 << 16);
    srv->gpu_us.reserve(1 << 16);
    srv->launch_log.reserve(1 << 16);
    srv->t_created = now_us();
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
        // results: pinned, mapped and coherent, so the kernel's stores reach host memory
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s.start) != hipSuccess || hipEventCreate(&s.done) != hipSuccess ||
            hipHostMalloc((void **)&s.h_in, in_bytes, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s.h_e, mb * kEStride, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer((void **)&s.h_e_dev, s.h_e, 0) != hipSuccess) {
            rc = set_error(CMD_ERR_HIP, "server slot allocation failed");
            break;
        }
        rc = dev_alloc(ctx, s.dev, per_slot);
        if (rc) break;
    }
    if (rc) {
        for (Slot &s : srv->slots) {
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.start) (void)hipEventDestroy(s.start);
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

extern "C" int cmd_server_submit_batch_mode(cmd_server *s, int32_t mode, size_t count, const int32_t *key_ids,
                                            const uint8_t *ct, size_t ct_stride, uint64_t *reqs, size_t *accepted)
{
    return submit_impl(s, mode, count, key_ids, ct, ct_stride, reqs, accepted);
}

extern "C" int cmd_server_submit_batch(cmd_server *s, size_t count, const int32_t *key_ids, const uint8_t *ct,
                                       size_t ct_stride, uint64_t *reqs, size_t *accepted)
{
    return submit_impl(s, CMD_MODE_FUSED, count, key_ids, ct, ct_stride, reqs, accepted);
}

extern "C" int cmd_server_submit(cmd_server *s, int32_t key_id, const uint8_t *ct, uint64_t *req)
{
    if (!s || !ct || !req) return CMD_ERR_ARG;
    return submit_impl(s, CMD_MODE_FUSED, 1, &key_id, ct, kMaxSyndBytes, req, nullptr);
}

extern "C" int cmd_server_result_bytes(cmd_server *s, uint64_t req, size_t *bytes)
{
    if (!s || !bytes) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const Req *r = s->lookup(req);
    if (!r) return set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)req);
    *bytes = r->out_bytes;
    return CMD_OK;
}

extern "C" int cmd_server_poll(cmd_server *s, uint64_t req, uint8_t *out, size_t out_cap)
{
    if (!s) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    return take_result(s, req, out, out_cap);
}

extern "C" int cmd_server_wait(cmd_server *s, uint64_t req, uint8_t *out, size_t out_cap, double timeout_us)
{
    if (!s) return CMD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const auto deadline = Clock::now() + std::chrono::microseconds((long long)(timeout_us < 0 ? 0 : timeout_us));
    for (;;) {
        int rc = take_result(s, req, out, out_cap);
        if (rc != CMD_ERR_PENDING) return rc;
        if (timeout_us < 0) {
            s->cv_done.wait(lk);
        } else if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            rc = take_result(s, req, out, out_cap);
            return rc == CMD_ERR_PENDING ? CMD_ERR_TIMEOUT : rc;
        }
    }
}

extern "C" int cmd_server_wait_batch(cmd_server *s, size_t count, const uint64_t *reqs, uint8_t *out,
                                     size_t out_stride, double timeout_us)
{
    if (!s || (count && (!reqs || !out))) return CMD_ERR_ARG;
    const auto deadline = Clock::now() + std::chrono::microseconds((long long)(timeout_us < 0 ? 0 : timeout_us));
    // The lock is held for at most kChunk requests at a time: the worker needs it to retire
    // launches, and a long hold here would stall every request in flight.
    constexpr size_t kChunk = 256;
    for (size_t i0 = 0; i0 < count; i0 += kChunk) {
        std::lock_guard<std::mutex> lk(s->mu);
        for (size_t i = i0; i < std::min(count, i0 + kChunk); i++) {
            const Req *r = s->lookup(reqs[i]);
            if (!r || r->state == kTaken)
                return set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)reqs[i]);
            if ((size_t)r->out_bytes > out_stride)
                return set_error(CMD_ERR_ARG, "out_stride %zu below the result of request %zu", out_stride, i);
        }
    }
    int first_err = CMD_OK;
    std::vector<uint32_t> taken;
    std::vector<size_t> pos;
    taken.reserve(kChunk);
    pos.reserve(kChunk);
    std::unique_lock<std::mutex> lk(s->mu);
    size_t i = 0;
    while (i < count) {
        // finished requests, in order, up to a chunk; wait only when nothing is taken yet
        taken.clear();
        pos.clear();
        int stop_rc = CMD_OK;
        while (i < count && taken.size() < kChunk) {
            Req *r = s->lookup(reqs[i]);
            if (!r || r->state == kTaken) {
                stop_rc = set_error(CMD_ERR_ARG, "unknown request %llu", (unsigned long long)reqs[i]);
                break;
            }
            if (r->state != kDone) {
                if (!taken.empty()) break;
                if (timeout_us < 0) {
                    s->cv_done.wait(lk);
                } else if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && r->state != kDone) {
                    stop_rc = CMD_ERR_TIMEOUT;
                    break;
                }
                continue;
            }
            r->state = kTaken;
            taken.push_back((uint32_t)reqs[i]);
            pos.push_back(i);
            i++;
        }
        lk.unlock();
        for (size_t k = 0; k < taken.size(); k++) {
            const Req &r = s->req(taken[k]);
            if (r.rc == CMD_OK)
                std::memcpy(out + pos[k] * out_stride, r.e, r.out_bytes);
            else if (first_err == CMD_OK)
                first_err = r.rc;
        }
        lk.lock();
        for (uint32_t idx : taken) {
            s->req(idx).state = kFree;
            s->free_idx.push_back(idx);
        }
        if (stop_rc != CMD_OK) return stop_rc;
    }
    return first_err;
}

extern "C" int cmd_server_stats_get(cmd_server *s, cmd_server_stats *out)
{
    if (!s || !out) return CMD_ERR_ARG;
    std::vector<double> lat, queue, dev, gpu;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        out->submitted = s->submitted;
        out->completed = s->completed;
        out->rejected = s->rejected;
        out->launches = s->launches;
        out->failed = s->failed;
        out->mean_batch = s->launches ? (double)s->launched_items / (double)s->launches : 0.0;
        double busy = s->busy_us;
        if (s->busy_since >= 0) busy += now_us() - s->busy_since;
        out->busy_s = busy * 1e-6;
        out->step_launch_max_us = s->step_max[0];
        out->step_retire_max_us = s->step_max[1];
        out->step_poll_max_us = s->step_max[2];
        out->wait_overshoot_max_us = s->step_max[3];
        lat = s->lat_us;
        queue = s->queue_us;
        dev = s->dev_us;
        gpu = s->gpu_us;
    }
    out->p50_us = percentile(lat, 0.50);
    out->p99_us = percentile(lat, 0.99);
    out->p999_us = percentile(lat, 0.999);
    out->max_us = percentile(lat, 1.0);
    out->queue_p99_us = percentile(queue, 0.99);
    out->queue_max_us = percentile(queue, 1.0);
    out->device_p99_us = percentile(dev, 0.99);
    out->device_max_us = percentile(dev, 1.0);
    out->gpu_p99_us = percentile(gpu, 0.99);
    out->gpu_max_us = percentile(gpu, 1.0);
    return CMD_OK;
}

extern "C" int cmd_server_latencies(cmd_server *s, double *total_us, double *queue_us, size_t cap, size_t *n)
{
    if (!s || (cap && (!total_us || !queue_us))) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const size_t k = std::min(cap, s->lat_us.size());
    if (k) {
        std::memcpy(total_us, s->lat_us.data(), k * sizeof(double));
        std::memcpy(queue_us, s->queue_us.data(), k * sizeof(double));
    }
    if (n) *n = s->lat_us.size();
    return CMD_OK;
}

extern "C" int cmd_server_launch_log(cmd_server *s, cmd_server_launch *out, size_t cap, size_t *n)
{
    if (!s || (cap && !out)) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const size_t k = std::min(cap, s->launch_log.size());
    if (k) std::memcpy(out, s->launch_log.data(), k * sizeof(cmd_server_launch));
    if (n) *n = s->launch_log.size();
    return CMD_OK;
}

extern "C" int cmd_server_reset_stats(cmd_server *s)
{
    if (!s) return CMD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->rejected = s->launches = s->launched_items = s->failed = 0;
    s->lat_us.clear();
    s->queue_us.clear();
    s->dev_us.clear();
    s->gpu_us.clear();
    s->launch_log.clear();
    for (double &m : s->step_max) m = 0;
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
        if (sl.start) (void)hipEventDestroy(sl.start);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        if (sl.h_e) (void)hipHostFree(sl.h_e);
        dev_free(s->ctx, sl.dev);
        dev_free(s->ctx, sl.kem);
    }
    delete s;  // every request was retired by the worker; the pool blocks go with the server
    return CMD_OK;
}
