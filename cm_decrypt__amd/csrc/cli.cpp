// cmdecrypt: drop-in driver for the reference's per-parameter decrypt executables.
//
//   cmdecrypt <param> [KATNUM] [--mode plain|fused] [--vectors DIR] [--out FILE]
//   decrypt_<param> [KATNUM]            (when invoked through a link of that name)
//
// Like CPU/param/param6960119/src/decrypt.c:109-218: KATNUM (default 5) must be a positive
// integer (else usage, exit -4); ct_<p>.bin / sk_<p>.bin are read from
// ../../../kem/test_vectors/Cipher_Sk relative to the working directory (util.c:12), and the
// errorstream goes to ../../results/output/errorstream0_<p>.bin (decrypt.c:18), reset first.
// A file error exits with -1.  The summary on stdout follows the GPU drivers' profile
// (GPU_Baseline/param/param6960119/src/cuda/decrypt.cu:74-114).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cm_decrypt.h"

static bool read_file(const std::string &path, std::vector<uint8_t> &buf, size_t need)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    buf.resize(need);
    size_t got = std::fread(buf.data(), 1, need, f);
    std::fclose(f);
    return got == need;
}

int main(int argc, char **argv)
{
    std::string prog = argv[0];
    size_t slash = prog.find_last_of('/');
    std::string base = slash == std::string::npos ? prog : prog.substr(slash + 1);
    int param = 0;
    int argi = 1;
    if (base.rfind("decrypt_", 0) == 0) {
        param = std::atoi(base.c_str() + 8);
    } else if (argc > 1) {
        param = std::atoi(argv[1]);
        argi = 2;
    }
    cmd_param_info info;
    if (cmd_param_get(param, &info) != CMD_OK) {
        std::fprintf(stderr, "Usage: %s <348864|460896|6688128|6960119|8192128> [KATNUM>0] [--mode plain|fused]\n",
                     argv[0]);
        return CMD_ERR_CRYPTO;
    }
    long katnum = 5;
    int mode = CMD_MODE_FUSED;
    std::string vec_dir = "../../../kem/test_vectors/Cipher_Sk";
    std::string out_path = "../../results/output/errorstream0_" + std::to_string(param) + ".bin";
    for (int i = argi; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--mode" && i + 1 < argc) {
            std::string m = argv[++i];
            mode = m == "plain" ? CMD_MODE_PLAIN : CMD_MODE_FUSED;
        } else if (a == "--vectors" && i + 1 < argc) {
            vec_dir = argv[++i];
        } else if (a == "--out" && i + 1 < argc) {
            out_path = argv[++i];
        } else {
            char *end = nullptr;
            katnum = std::strtol(a.c_str(), &end, 10);
            if (a.empty() || *end != '\0' || katnum <= 0) {
                std::fprintf(stderr, "Usage: %s [KATNUM>0]\n", argv[0]);
                return CMD_ERR_CRYPTO;
            }
        }
    }

    std::vector<uint8_t> sk;
    const std::string p = std::to_string(param);
    FILE *fct = nullptr;
    if (!read_file(vec_dir + "/sk_" + p + ".bin", sk, (size_t)info.sk_bytes) ||
        !(fct = std::fopen((vec_dir + "/ct_" + p + ".bin").c_str(), "rb"))) {
        std::perror("failed to read test vectors");
        return CMD_ERR_FILE;
    }
    // a short ciphertext file fails before the device is touched, like the reference's fread
    // the fused path brings back error positions (2t bytes per ciphertext) instead of the
    // n/8-byte bitmaps; the plain path keeps the bitmaps of the per-stage kernels
    const size_t row = mode == CMD_MODE_PLAIN ? (size_t)info.e_bytes : (size_t)2 * info.t;
    const size_t ct_bytes = (size_t)katnum * info.ct_bytes, e_bytes = (size_t)katnum * row;
    if (std::fseek(fct, 0, SEEK_END) != 0 || std::ftell(fct) < (long)ct_bytes || std::fseek(fct, 0, SEEK_SET) != 0) {
        std::fprintf(stderr, "failed to read test vectors: ct_%s.bin holds fewer than %ld ciphertexts\n", p.c_str(),
                     katnum);
        std::fclose(fct);
        return CMD_ERR_FILE;
    }

    cmd_ctx *ctx = nullptr;
    int rc = cmd_ctx_create(0, 0, &ctx);
    if (rc) {
        std::fprintf(stderr, "cmd_ctx_create: %s (%s)\n", cmd_strerror(rc), cmd_last_error());
        std::fclose(fct);
        return rc;
    }
    // ciphertexts are read straight into pinned memory and the error vectors come back into
    // pinned memory, so both cross PCIe by DMA
    void *ct = nullptr, *e = nullptr;
    rc = cmd_host_alloc(ctx, ct_bytes, &ct);
    if (!rc) rc = cmd_host_alloc(ctx, e_bytes, &e);
    if (rc) {
        std::fprintf(stderr, "cmd_host_alloc: %s (%s)\n", cmd_strerror(rc), cmd_last_error());
        std::fclose(fct);
        cmd_ctx_destroy(ctx);
        return rc;
    }
    const size_t got = std::fread(ct, 1, ct_bytes, fct);
    std::fclose(fct);
    if (got != ct_bytes) {
        std::fprintf(stderr, "failed to read test vectors: %s holds fewer than %ld ciphertexts\n",
                     (vec_dir + "/ct_" + p + ".bin").c_str(), katnum);
        cmd_ctx_destroy(ctx);
        return CMD_ERR_FILE;
    }
    int32_t key = -1;
    rc = cmd_key_load(ctx, param, sk.data(), sk.size(), &key);
    if (rc) {
        std::fprintf(stderr, "cmd_key_load: %s (%s)\n", cmd_strerror(rc), cmd_last_error());
        cmd_ctx_destroy(ctx);
        return rc;
    }
    std::vector<int32_t> keys((size_t)katnum, key);
    auto t0 = std::chrono::steady_clock::now();
    rc = cmd_decrypt(ctx, mode == CMD_MODE_PLAIN ? CMD_MODE_PLAIN : CMD_MODE_POSITIONS, (size_t)katnum, keys.data(),
                     (const uint8_t *)ct, (size_t)info.ct_bytes, (uint8_t *)e, row);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) {
        std::fprintf(stderr, "cmd_decrypt: %s (%s)\n", cmd_strerror(rc), cmd_last_error());
        cmd_ctx_destroy(ctx);
        return rc;
    }
    rc = mode == CMD_MODE_PLAIN
             ? cmd_errorstream_write(param, out_path.c_str(), (const uint8_t *)e, row, (size_t)katnum, 0)
             : cmd_errorstream_write_positions(param, out_path.c_str(), (const uint16_t *)e, row, (size_t)katnum, 0);
    cmd_ctx_destroy(ctx);
    if (rc) {
        std::fprintf(stderr, "failed to write %s\n", out_path.c_str());
        return CMD_ERR_FILE;
    }
    const double wall = std::chrono::duration<double>(t1 - t0).count();
    std::printf("===== MI355X %s pipeline, mceliece%d =====\n", mode == CMD_MODE_PLAIN ? "plain" : "fused", param);
    std::printf("ciphertexts processed : %ld\n", katnum);
    std::printf("wall (H2D+kernels+D2H): %.3f ms\n", wall * 1e3);
    std::printf("throughput            : %.2f ct/s\n", (double)katnum / wall);
    std::printf("errorstream           : %s\n", out_path.c_str());
    return 0;
}
