// emu_driver.cpp — runs the kernel sources on the CPU (hip_emu.hpp) through the same C ABI as
// the device library. Test infrastructure: built with AddressSanitizer by the CPU test suite.
//
// usage: isar_emu_driver <input.bin> <output.bin> [autofocus_mode [ra_split [static [czt_length [czt_threads]]]]]
// static = 1: also images the batch, then its first frame alone, through isar_run_device_static
// and requires both to equal the context's images bit for bit (exit 10-12 otherwise).
// input : int32 M, N, W, K; int64 band_width, fc; double Tp; int64 Fs; isar_options;
//         echo K*M*N complex float; velocity K*M double; [phi0 K*M complex float if warm_start]
// output: img K*M*W float; then per frame: int32[10] scalars, float[2], double lag[M-1], imax[M-1],
//         colmax[W], colE[W], colH[W], select_bin[W], phi0[2M], phi[2M]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "isar_imaging_ext.h"

int main(int argc, char** argv) {
    if (argc < 3 || argc > 8) {
        std::fprintf(stderr, "usage: %s input.bin output.bin [autofocus_mode [ra_split [static [czt_length [czt_threads]]]]]\n",
                     argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    int32_t hdr[4];
    int64_t B, fc, Fs;
    double Tp;
    isar_options opt;
    bool ok = std::fread(hdr, sizeof(hdr), 1, in) == 1 && std::fread(&B, 8, 1, in) == 1 && std::fread(&fc, 8, 1, in) == 1 &&
              std::fread(&Tp, 8, 1, in) == 1 && std::fread(&Fs, 8, 1, in) == 1 && std::fread(&opt, sizeof(opt), 1, in) == 1;
    if (!ok) return 4;
    const int M = hdr[0], N = hdr[1], W = hdr[2], K = hdr[3];
    std::vector<float> echo(2ull * K * M * N), phi0(2ull * K * M);
    std::vector<double> vel(static_cast<size_t>(K) * M);
    ok = std::fread(echo.data(), sizeof(float), echo.size(), in) == echo.size() &&
         std::fread(vel.data(), sizeof(double), vel.size(), in) == vel.size();
    if (ok && opt.warm_start) ok = std::fread(phi0.data(), sizeof(float), phi0.size(), in) == phi0.size();
    std::fclose(in);
    if (!ok) return 5;
    isar_radar_params p{M, N, W, 0, B, fc, Tp, Fs};
    isar_context* ctx = nullptr;
    if (isar_create(&p, K, &ctx) != 0) {
        std::fprintf(stderr, "create: %s\n", isar_last_error());
        return 6;
    }
    if (argc >= 5 && isar_set_ra_split(ctx, std::atoi(argv[4])) != 0) {
        std::fprintf(stderr, "ra split: %s\n", isar_last_error());
        return 6;
    }
    if (argc >= 4 && isar_set_autofocus_mode(ctx, std::atoi(argv[3])) != 0) {
        std::fprintf(stderr, "autofocus mode: %s\n", isar_last_error());
        return 6;
    }
    if (argc >= 7 && isar_set_czt_length(ctx, std::atoi(argv[6])) != 0) {
        std::fprintf(stderr, "chirp-z length: %s\n", isar_last_error());
        return 6;
    }
    if (argc >= 8 && isar_set_czt_threads(ctx, std::atoi(argv[7])) != 0) {
        std::fprintf(stderr, "chirp-z threads: %s\n", isar_last_error());
        return 6;
    }
    std::vector<float> img(static_cast<size_t>(K) * M * W);
    if (isar_run_host(ctx, K, echo.data(), vel.data(), opt.warm_start ? phi0.data() : nullptr, img.data(), &opt) != 0) {
        std::fprintf(stderr, "run: %s\n", isar_last_error());
        return 7;
    }
    if (argc >= 6 && std::atoi(argv[5]) == 1) {
        // the emulated device memory is host memory: the host arrays serve as device buffers
        std::vector<float> img_s(img.size());
        const void* ph = opt.warm_start ? phi0.data() : nullptr;
        if (isar_run_device_static(&p, K, echo.data(), vel.data(), ph, img_s.data(), &opt, nullptr) != 0) {
            std::fprintf(stderr, "static run: %s\n", isar_last_error());
            return 10;
        }
        if (img_s != img) return 11;
        std::vector<float> one(static_cast<size_t>(M) * W);
        if (isar_run_device_static(&p, 1, echo.data(), vel.data(), ph, one.data(), &opt, nullptr) != 0 ||
            !std::equal(one.begin(), one.end(), img.begin()))
            return 12;
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 8;
    std::fwrite(img.data(), sizeof(float), img.size(), out);
    for (int f = 0; f < K; ++f) {
        std::vector<double> lag(M - 1);
        std::vector<float> cm(W), ce(W), ch(W), p0(2 * M), p1(2 * M);
        std::vector<int32_t> imax(M - 1), sel(W, -1);
        isar_frame_trace t{};
        t.ra_lag = lag.data();
        t.ra_imax = imax.data();
        t.pa_colmax = cm.data();
        t.pa_colE = ce.data();
        t.pa_colH = ch.data();
        t.select_bin = sel.data();
        t.phi0 = p0.data();
        t.phi = p1.data();
        if (isar_get_trace(ctx, f, &t) != 0) return 9;
        int32_t sc[10] = {t.n_pairs, t.center_applied, t.center_shift, t.idx1_len, t.idx2_len,
                          t.diff_imin, t.cut_offset, t.tgt_num, t.unit2_num, t.status};
        float fs[2] = {t.center_thr, t.pa_threshold};
        std::fwrite(sc, sizeof(sc), 1, out);
        std::fwrite(fs, sizeof(fs), 1, out);
        std::fwrite(lag.data(), sizeof(double), lag.size(), out);
        std::fwrite(imax.data(), sizeof(int32_t), imax.size(), out);
        std::fwrite(cm.data(), sizeof(float), W, out);
        std::fwrite(ce.data(), sizeof(float), W, out);
        std::fwrite(ch.data(), sizeof(float), W, out);
        std::fwrite(sel.data(), sizeof(int32_t), W, out);
        std::fwrite(p0.data(), sizeof(float), p0.size(), out);
        std::fwrite(p1.data(), sizeof(float), p1.size(), out);
    }
    std::fclose(out);
    isar_destroy(ctx);
    return 0;
}
