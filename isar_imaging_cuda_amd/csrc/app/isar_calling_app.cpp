// isar_calling_app.cpp — the reference's sample caller (calling_dll/calling_dll.cpp:17-89) against
// the drop-in library: gpuDevInit -> dataParsing -> imagingMemInit -> { dataExtracting ->
// isarMainSingle } x n -> imagingMemDest. Built against include/isar_imaging_export.h only.
//
// usage: isar_calling_app <dir> <window_len> <imaging_stride> <n_images> <out_prefix> [polar] [mtrc] [hpc] [data_type]
//                         [hrrp] [dump_dir] [af_wait_polls]  (defaults: polar_type 1, if_mtrc 1, if_hpc 1,
//                         data_type 2, if_hrrp 0, no dump, the library's wait bound)
// writes <out_prefix>_<i>.bin (raw float32 image, window_len x 512) and <out_prefix>_0.txt; with
// if_hrrp also <out_prefix>_hrrp_<i>.bin (window_len x range_num complex float range profiles);
// with dump_dir the first window's intermediate arrays as the reference's text dumps
// (dataW.dat ... final.dat).
// af_wait_polls sets the session context's resident-autofocus wait bound (isar_set_af_wait_polls;
// -1 makes every wait give up, so each window is imaged again with the split form).
// n_images = 0 only parses the directory (no GPU call) and prints the decoded headers.
// n_images = -1 checks that dataExtracting refuses arguments that do not match imagingMemInit
// (frame length, data type, a too-small output vector) instead of overrunning a buffer.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "isar_imaging_export.h"
#include "isar_imaging_ext.h"

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s dir window_len imaging_stride n_images out_prefix [polar] [mtrc] [hpc] [data_type] [hrrp] [dump_dir]\n",
                     argv[0]);
        return 2;
    }
    const std::string dir_path = argv[1];
    int window_len = std::atoi(argv[2]);
    int imaging_stride = std::atoi(argv[3]);
    const int n_images = std::atoi(argv[4]);
    const std::string prefix = argv[5];
    const int polar_type = argc > 6 ? std::atoi(argv[6]) : 1;
    const bool if_mtrc = argc > 7 ? std::atoi(argv[7]) != 0 : true;
    const bool if_hpc = argc > 8 ? std::atoi(argv[8]) != 0 : true;
    const int data_type = argc > 9 ? std::atoi(argv[9]) : 2;   // 1 IFDS, 2 STRETCH
    const int sampling_stride = 1;
    int window_head = 0;
    const int option_alignment = 0, option_phase = 1;
    const bool if_hrrp = argc > 10 ? std::atoi(argv[10]) != 0 : false;
    const std::string dump_dir = argc > 11 ? argv[11] : "";
    const bool set_polls = argc > 12;
    const int af_wait_polls = set_polls ? std::atoi(argv[12]) : 0;

    vec1D_DBL dataN;
    vec1D_FLT turnAngle;
    int frame_len = 0, frame_num = 0;
    vec1D_INT dataWFileSn;
    vec1D_DBL dataNOut;
    vec1D_FLT turnAngleOut;
    vec1D_COM_FLT dataW;
    vec1D_FLT img;

    if (n_images != 0 && gpuDevInit() != EXIT_SUCCESS) {
        std::fprintf(stderr, "gpuDevInit: %s\n", isar_last_error());
        return 3;
    }
    if (dataParsing(&dataN, &turnAngle, &frame_len, &frame_num, dir_path, polar_type, data_type) != EXIT_SUCCESS) {
        std::fprintf(stderr, "dataParsing: %s\n", isar_last_error());
        return 4;
    }
    if (n_images == 0) {
        const size_t total = dataN.size() / 4;
        std::printf("frame_len %d frame_num %d frames %zu\n", frame_len, frame_num, total);
        for (size_t i = 0; i < total; ++i)
            std::printf("%zu %.6f %.6f %.9f %.9f %.7f\n", i, dataN[i], dataN[i + total], dataN[i + 2 * total],
                        dataN[i + 3 * total], turnAngle[i]);
        return 0;
    }
    imagingMemInit(&img, &dataWFileSn, &dataNOut, &turnAngleOut, &dataW, window_len, frame_len, data_type, if_hpc, if_hrrp);
    if (img.size() != static_cast<size_t>(window_len) * 512) {
        std::fprintf(stderr, "imagingMemInit: %s\n", isar_last_error());
        return 5;
    }
    if (set_polls && isar_set_af_wait_polls(isar_session_context(), af_wait_polls) != ISAR_OK) {
        std::fprintf(stderr, "isar_set_af_wait_polls: %s\n", isar_last_error());
        return 5;
    }
    if (n_images < 0) {
        int refused = 0;
        auto check = [&](const char* what, int fl, int dt, vec1D_COM_FLT* w) {
            const int rc = dataExtracting(&dataWFileSn, &dataNOut, &turnAngleOut, w, dataN, turnAngle, fl, frame_num,
                                          sampling_stride, window_head, window_len, imaging_stride, dt);
            std::printf("%s: rc %d (%s)\n", what, rc, rc == EXIT_SUCCESS ? "" : isar_last_error());
            refused += rc != EXIT_SUCCESS;
        };
        vec1D_COM_FLT small(dataW.size() / 2);
        check("frame_len + 4", frame_len + 4, data_type, &dataW);
        check("other data_type", frame_len, data_type == 2 ? 1 : 2, &dataW);
        if (data_type == 2) check("dataW half size", frame_len, data_type, &small);
        else ++refused;
        check("matching", frame_len, data_type, &dataW);   // must still work afterwards
        imagingMemDest(data_type, if_hpc, if_hrrp);
        return refused == 3 ? 0 : 9;
    }
    for (int i = 0; i < n_images; ++i) {
        auto t1 = std::chrono::high_resolution_clock::now();
        if (dataExtracting(&dataWFileSn, &dataNOut, &turnAngleOut, &dataW, dataN, turnAngle, frame_len, frame_num,
                           sampling_stride, window_head, window_len, imaging_stride, data_type) != EXIT_SUCCESS) {
            std::fprintf(stderr, "dataExtracting: %s\n", isar_last_error());
            return 6;
        }
        auto t2 = std::chrono::high_resolution_clock::now();
        isarMainSingle(img.data(), data_type, option_alignment, option_phase, if_hrrp, if_hpc, if_mtrc);
        auto t3 = std::chrono::high_resolution_clock::now();
        std::printf("[img %2d] %3dms / %3dms\n", i + 1,
                    static_cast<int>(std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t1).count()),
                    static_cast<int>(std::chrono::duration_cast<std::chrono::milliseconds>(t3 - t2).count()));
        std::ofstream out(prefix + "_" + std::to_string(i) + ".bin", std::ios::binary);
        out.write(reinterpret_cast<const char*>(img.data()), static_cast<std::streamsize>(img.size() * sizeof(float)));
        if (i == 0) writeFileFLT(prefix + "_0.txt", img.data(), img.size());
        if (i == 0 && !dump_dir.empty() && isar_session_dump((1u << ISAR_DUMP_COUNT) - 1, dump_dir.c_str()) != ISAR_OK) {
            std::fprintf(stderr, "isar_session_dump: %s\n", isar_last_error());
            return 8;
        }
        if (if_hrrp) {
            vec1D_FLT hrrp(img.size() / 512 * static_cast<size_t>(data_type == 1 ? 4096 : (frame_len - 256) / 4) * 2);
            if (isar_session_hrrp(hrrp.data()) != ISAR_OK) {
                std::fprintf(stderr, "isar_session_hrrp: %s\n", isar_last_error());
                return 7;
            }
            std::ofstream h(prefix + "_hrrp_" + std::to_string(i) + ".bin", std::ios::binary);
            h.write(reinterpret_cast<const char*>(hrrp.data()), static_cast<std::streamsize>(hrrp.size() * sizeof(float)));
        }
        window_head += imaging_stride;
    }
    imagingMemDest(data_type, if_hpc, if_hrrp);
    std::printf("turn angle of the last window: %.6f .. %.6f\n", turnAngleOut.front(), turnAngleOut.back());
    return 0;
}
