// Host engine internals shared by the C ABI, the batch planner and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cm_decrypt.h"
#include "kernels.h"
#include "keys.h"
#include "params.h"

namespace cmd {

int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define CMD_HIP(call)                                   \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return hip_fail(_e, #call); \
    } while (0)

// Device allocation charged against a context budget.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

constexpr size_t kMaxKeys = 16384;  // resident key slots per context

struct KeySlot {
    bool used = false;
    int pidx = -1;
    int inflight = 0;      // server requests and live batches referencing the key (context lock)
    KeyDev dev{};
    void *mem = nullptr;   // one allocation holding table, L, lbits
    size_t bytes = 0;
};

struct Engine;

// One stage of the host-buffer pipeline (hostio.cpp): the device buffers of one chunk, pinned
// staging, the chunk's plan and the events that order it between the pipeline's copy and
// kernel streams.
constexpr size_t kIoSlots = 3;
struct IoSlot {
    // in_done: copy in finished; k_done: kernels finished; done: copy out finished;
    // t0/t1 bracket the kernels when timing
    hipEvent_t in_done = nullptr, k_done = nullptr, done = nullptr, t0 = nullptr, t1 = nullptr;
    uint8_t *h_ct = nullptr, *h_out = nullptr;
    char *h_plan = nullptr;
    size_t h_ct_cap = 0, h_out_cap = 0, h_plan_cap = 0;
    DevBuf d_ct, d_out, d_plan, d_kem_e, d_ok, d_msg;
    bool busy = false, staged_out = false, timed = false;
    bool out_pending = false;   // kernels enqueued, copy out not yet enqueued
    bool mark_pending = false;  // copy out enqueued, `done` not yet recorded
    uint8_t *dst = nullptr;    // where the copy out goes (caller's buffer or h_out)
    size_t first = 0, count = 0, out_bytes = 0;
    // the chunk's plan: schedules, tiles, and the offsets of its parts in the plan image
    FusedSchedule sched[kNumParams];
    std::vector<FusedTile> tiles;
    int launch_tile = 32;
    size_t off_keys = 0, off_tiles = 0, off_ctr = 0;
};

}  // namespace cmd

// Opaque handles of the C ABI.
struct cmd_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, used = 0;
    std::recursive_mutex lock;

    std::vector<cmd::KeySlot> keys;
    cmd::KeyDev *d_keys = nullptr;  // kMaxKeys descriptors, slot i mirrors keys[i].dev

    // work buffers of cmd_decrypt / cmd_decrypt_device (grow on demand)
    cmd::DevBuf ct, e, s, loc, idx, plan;
    cmd::DevBuf keystage;  // secret-key bytes for on-device expansion
    uint32_t *d_counter = nullptr;

    cmd::FusedSchedule sched[cmd::kNumParams];
    cmd::FftTab *d_fft = nullptr;   // kFftVariants additive-FFT constant tables

    std::vector<cmd::IoSlot> io;                            // host-buffer pipeline stages
    hipStream_t io_copy = nullptr, io_kern = nullptr;       // the pipeline's copies; its kernels
    std::vector<std::pair<uintptr_t, size_t>> host_pinned;  // cmd_host_alloc blocks

    // host plan of cmd_decrypt_device_noalloc, reused from call to call
    std::vector<cmd::FusedTile> na_tiles;
    std::vector<int32_t> na_items, na_item_key;

    bool timing = false;
    bool fit_tiles = true;  // batches narrow their tiles when too small to fill the GPU (off while autotuning)
    double timing_us = 0.0;
    uint64_t timing_n = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct cmd_batch {
    cmd_ctx *ctx = nullptr;
    int mode = 0;
    size_t count = 0;
    size_t ct_stride = 0, e_stride = 0;
    bool use_graph = false;

    // plain: groups of one key
    struct Group { int pidx; int key; int first; int count; };
    std::vector<Group> groups;
    int max_group = 0;
    int max_t = 0;

    // fused: tile list and per-slot item/key arrays
    std::vector<cmd::FusedTile> tiles;
    std::vector<int32_t> items, item_key;
    int launch_tile = 32;      // kernel tile width: the widest tile of the batch
    std::vector<int32_t> pinned;  // distinct keys of the batch, held against unload
    cmd::DevBuf d_plan;       // device copy of idx / tiles / items / keys
    int32_t *d_idx = nullptr; // plain: item order (grouped)
    cmd::FusedTile *d_tiles = nullptr;
    int32_t *d_items = nullptr, *d_item_key = nullptr;
    cmd::DevBuf d_s, d_loc;   // plain intermediates
    cmd::DevBuf d_kem_e, d_ok, d_msg;  // KEM: decoded error vectors, accept flags, SHAKE256 inputs
    size_t kem_e_stride = 0;
    uint32_t *d_counter = nullptr;
    cmd::FusedSchedule sched[cmd::kNumParams];

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graph_ct = nullptr;
    void *graph_e = nullptr;
    hipStream_t graph_stream = nullptr;
    hipEvent_t done = nullptr;  // recorded after each run on its stream; destroy waits for it
    bool ran = false;
};

namespace cmd {

int dev_alloc(cmd_ctx *ctx, DevBuf &b, size_t bytes);    // exact size, charged
int dev_reserve(cmd_ctx *ctx, DevBuf &b, size_t bytes);  // grow only
void dev_free(cmd_ctx *ctx, DevBuf &b);

// Host plan of a fused launch over `count` items: tiles of one parameter set each (sorted
// by set, then key, largest sets first) with per-slot item index and key.  Returns the
// widest tile in *launch_tile.
int plan_fused(const cmd_ctx *ctx, const FusedSchedule sched[], size_t count, const int32_t *key_ids,
               std::vector<FusedTile> &tiles, std::vector<int32_t> &items, std::vector<int32_t> &item_key,
               int *launch_tile);
// Schedule a fused launch uses: tile width of the batch, waves/groups of its heaviest set.
FusedSchedule launch_schedule(const FusedSchedule sched[], const std::vector<FusedTile> &tiles, int launch_tile);

// Schedules for a batch of `count` items: the tuned ones, with every tile width above
// kMinFitTile halved until the batch has about four tiles per CU (or cannot narrow further).
// A tile's Berlekamp-Massey costs the same for 4 or 32 ciphertexts, so a batch too small to
// fill the GPU at the tuned width finishes sooner in more, narrower tiles.
constexpr int kMinFitTile = 4;
void fit_tiles(const cmd_ctx *ctx, const FusedSchedule in[], size_t count, const int32_t *key_ids,
               FusedSchedule out[]);
// Output row bytes a mode needs for these items (keys already validated).
size_t out_row_bytes(int mode, size_t count, const int32_t *key_ids, const cmd_ctx *ctx);
// Builds the plan of a batch (host arrays) and uploads it.
int batch_plan(cmd_batch *b, const int32_t *key_ids);
// Enqueues the batch's launches on `st` (no graph).
int batch_enqueue(cmd_batch *b, const void *d_ct, void *d_e, hipStream_t st);

// Host-buffer decryption (fused or KEM mode) through the chunked pipeline; io_release frees
// its slots and the context's pinned blocks.
int io_decrypt(cmd_ctx *ctx, int mode, size_t count, const int32_t *key_ids, const uint8_t *ct, size_t ct_stride,
               uint8_t *out, size_t out_stride);
void io_release(cmd_ctx *ctx);

// Device scratch of a KEM launch: decoded error vectors, accept flags, SHAKE256 inputs.
struct KemScratch {
    uint8_t *e;
    size_t e_stride;
    uint32_t *ok;
    uint8_t *msg;
};
// Enqueues a planned fused launch (tiles already on the device, d_counter one word) on `st`;
// for CMD_MODE_KEM also the check and the session-key kernels, `out` receiving the keys.
int enqueue_fused(cmd_ctx *ctx, int mode, const FusedSchedule sched[], const std::vector<FusedTile> &tiles,
                  int launch_tile, const FusedTile *d_tiles, const int32_t *d_items, const int32_t *d_item_key,
                  uint32_t *d_counter, const uint8_t *ct, size_t ct_stride, uint8_t *out, size_t out_stride,
                  const KemScratch &kem, hipStream_t st);
void batch_release(cmd_batch *b);

FusedSchedule default_schedule(int pidx);

// Looks up this device's copy of the allocation-free entry point's device array (once per
// device; the device must be current).
int noalloc_setup(int device);

}  // namespace cmd
