// Schedule autotuner for the fused kernel and JSON persistence of the tuned schedules.
//
// The decoder's work is independent of the data (constant-time structure), so random
// ciphertext bytes time exactly like real ones; candidates are timed on the device with
// HIP events and the fastest schedule per parameter set is kept.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "engine.h"

using namespace cmd;

static std::vector<FusedSchedule> candidates()
{
    std::vector<FusedSchedule> out;
    for (int alg : {1, 0})
        for (int tile : {4, 8, 16, 32})
            for (int wpg : {1, 2, 4})
                for (int gpc : {1, 2, 4, 8}) {
                    FusedSchedule s{tile, wpg, 1, gpc, alg};
                    // LDS per workgroup is wpg * 16.4 KB; keep resident groups within 160 KB per CU
                    if (wpg * gpc > 8) continue;  // at most two wavefronts per SIMD
                    if (fused_supported(s)) out.push_back(s);
                }
    // two wavefronts per tile (one pair per workgroup, 128-register build): up to eight pairs,
    // sixteen wavefronts, per CU; a pair's LDS is 17 KB
    for (int tile : {8, 16, 32})
        for (int gpc : {2, 3, 4, 5, 6, 7, 8}) {
            FusedSchedule s{tile, 2, 2, gpc, 1};
            if (fused_supported(s)) out.push_back(s);
        }
    return out;
}

extern "C" int cmd_autotune(cmd_ctx *ctx, int32_t param_id, size_t count, const int32_t *key_ids, int32_t reps,
                            cmd_schedule *best, double *best_us)
{
    if (!ctx || !key_ids || !count) return CMD_ERR_ARG;
    int pidx = param_index(param_id);
    if (pidx < 0) return set_error(CMD_ERR_PARAM, "unknown parameter set %d", (int)param_id);
    if (reps < 1) reps = 3;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    CMD_HIP(hipSetDevice(ctx->device));
    for (size_t i = 0; i < count; i++) {
        int32_t k = key_ids[i];
        if (k < 0 || (size_t)k >= ctx->keys.size() || !ctx->keys[k].used || ctx->keys[k].pidx != pidx)
            return set_error(CMD_ERR_KEY, "autotune item %zu: key %d is not a loaded key of set %d", i, (int)k,
                             (int)param_id);
    }
    const ParamDesc &p = kParams[pidx];
    const size_t ct_stride = (size_t)p.synd_bytes, e_stride = (size_t)((p.n / 8 + 3) / 4 * 4);
    DevBuf dct, de;
    int rc = dev_alloc(ctx, dct, count * ct_stride);
    if (!rc) rc = dev_alloc(ctx, de, count * e_stride);
    if (rc) {
        dev_free(ctx, dct);
        return rc;
    }
    std::vector<uint8_t> host(count * ct_stride);
    std::mt19937_64 rng(0x5EEDu + (uint64_t)param_id);
    for (auto &b : host) b = (uint8_t)rng();
    hipError_t he = hipMemcpy(dct.ptr, host.data(), host.size(), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        dev_free(ctx, dct);
        dev_free(ctx, de);
        return hip_fail(he, "autotune upload");
    }

    const FusedSchedule saved = ctx->sched[pidx];
    FusedSchedule winner = saved;
    const bool fit_saved = ctx->fit_tiles;
    ctx->fit_tiles = false;  // time each candidate's own tile width
    double winner_us = 1e30;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (const FusedSchedule &s : candidates()) {
        ctx->sched[pidx] = s;
        cmd_batch *b = nullptr;
        rc = cmd_batch_create(ctx, CMD_MODE_FUSED, count, key_ids, ct_stride, e_stride, 0, &b);
        if (rc) break;
        double best_here = 1e30;
        for (int r = 0; r < reps + 1 && !rc; r++) {
            (void)hipEventRecord(e0, ctx->stream);
            rc = cmd_batch_run(b, dct.ptr, de.ptr, ctx->stream);
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0) best_here = std::min(best_here, (double)ms * 1000.0);  // first run warms up
        }
        cmd_batch_destroy(b);
        if (rc) break;
        if (best_here < winner_us) {
            winner_us = best_here;
            winner = s;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(ctx, dct);
    dev_free(ctx, de);
    ctx->fit_tiles = fit_saved;
    ctx->sched[pidx] = rc ? saved : winner;
    if (rc) return rc;
    if (best) {
        best->tile_cts = winner.tile_cts;
        best->waves_per_group = winner.waves_per_group;
        best->waves_per_ct = winner.waves_per_ct;
        best->groups_per_cu = winner.groups_per_cu;
        best->root_alg = winner.root_alg;
    }
    if (best_us) *best_us = winner_us;
    return CMD_OK;
}

extern "C" int cmd_schedule_save(cmd_ctx *ctx, const char *path)
{
    if (!ctx || !path) return CMD_ERR_ARG;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    FILE *f = std::fopen(path, "w");
    if (!f) return set_error(CMD_ERR_FILE, "cannot write %s", path);
    std::fprintf(f, "{\n");
    for (int i = 0; i < kNumParams; i++) {
        const FusedSchedule &s = ctx->sched[i];
        std::fprintf(f, "  \"%d\": {\"tile_cts\": %d, \"waves_per_group\": %d, \"waves_per_ct\": %d, "
                        "\"groups_per_cu\": %d, \"root_alg\": %d}%s\n",
                     kParams[i].id, s.tile_cts, s.waves_per_group, s.waves_per_ct, s.groups_per_cu, s.root_alg,
                     i + 1 < kNumParams ? "," : "");
    }
    std::fprintf(f, "}\n");
    return std::fclose(f) == 0 ? CMD_OK : set_error(CMD_ERR_FILE, "cannot write %s", path);
}

// Reads the JSON written by cmd_schedule_save (or any layout of the same object: per set id,
// an object with the five fields in any order and spacing).
static bool json_int_field(const std::string &obj, const char *name, int *out)
{
    const std::string key = std::string("\"") + name + "\"";
    size_t at = obj.find(key);
    if (at == std::string::npos) return false;
    at = obj.find(':', at + key.size());
    if (at == std::string::npos) return false;
    char *end = nullptr;
    const long v = std::strtol(obj.c_str() + at + 1, &end, 10);
    if (end == obj.c_str() + at + 1) return false;
    *out = (int)v;
    return true;
}

extern "C" int cmd_schedule_load(cmd_ctx *ctx, const char *path)
{
    if (!ctx || !path) return CMD_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return set_error(CMD_ERR_FILE, "cannot read %s", path);
    std::string text;
    char buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, n);
    std::fclose(f);
    int loaded = 0;
    std::lock_guard<std::recursive_mutex> g(ctx->lock);
    for (int pidx = 0; pidx < kNumParams; pidx++) {
        const std::string key = "\"" + std::to_string(kParams[pidx].id) + "\"";
        const size_t at = text.find(key);
        if (at == std::string::npos) continue;
        const size_t open = text.find('{', at), close = text.find('}', at);
        if (open == std::string::npos || close == std::string::npos || close < open) continue;
        const std::string obj = text.substr(open, close - open + 1);
        FusedSchedule s{};
        if (!json_int_field(obj, "tile_cts", &s.tile_cts) || !json_int_field(obj, "waves_per_group", &s.waves_per_group) ||
            !json_int_field(obj, "waves_per_ct", &s.waves_per_ct) || !json_int_field(obj, "groups_per_cu", &s.groups_per_cu) ||
            !json_int_field(obj, "root_alg", &s.root_alg))
            continue;
        if (fused_supported(s) && s.groups_per_cu >= 1 && s.groups_per_cu <= 16) {
            ctx->sched[pidx] = s;
            loaded++;
        }
    }
    return loaded ? CMD_OK : set_error(CMD_ERR_FILE, "no schedules in %s", path);
}
