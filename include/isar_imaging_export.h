/*
 * isar_imaging_export.h — drop-in replacement of the reference's exported API
 * (isar_imaging/isar_imaging_export.h:37-130), same names, argument types and call order.
 *
 * The seven entry points keep C linkage (unmangled names) and the reference's C++ parameter
 * types (std::vector / std::string references), so a caller built against the reference header
 * links against libisar_imaging.so unchanged (same libstdc++ ABI required, as for the original).
 * Differences a caller can observe:
 *   - nothing calls exit(): int functions return EXIT_FAILURE and void functions leave an error
 *     message in isar_last_error() (isar_imaging_ext.h) instead of terminating;
 *   - dataExtracting checks its arguments against imagingMemInit and dataParsing (data type,
 *     frame length, window length, vector sizes) and returns EXIT_FAILURE on a mismatch, where
 *     the reference would overrun a buffer. Both ingests are provided: STRETCH dechirped frames
 *     and IFDS raw echoes, whose pulse compression runs on the GPU.
 */
#ifndef ISAR_IMAGING_EXPORT_H_
#define ISAR_IMAGING_EXPORT_H_

#include <complex>
#include <string>
#include <vector>

#if defined(_WIN32)
#define DLL_EXPORT_API __declspec(dllexport)
#else
#define DLL_EXPORT_API __attribute__((visibility("default")))
#endif

typedef std::vector<int> vec1D_INT;
typedef std::vector<std::vector<int>> vec2D_INT;
typedef std::vector<float> vec1D_FLT;
typedef std::vector<std::vector<float>> vec2D_FLT;
typedef std::vector<double> vec1D_DBL;
typedef std::vector<std::vector<double>> vec2D_DBL;
typedef std::vector<std::complex<float>> vec1D_COM_FLT;
typedef std::vector<std::vector<std::complex<float>>> vec2D_COM_FLT;

/* Selects the GPU (isar_imaging_export.h:37; reference: findCudaDevice). */
extern "C" DLL_EXPORT_API int gpuDevInit();

/* Reads the .wbd directory: system header, per-frame navigation data, turn-angle curve
 * (isar_imaging_export.h:52; reference body isar_imaging_export.cpp:48-70). */
extern "C" DLL_EXPORT_API int dataParsing(vec1D_DBL* dataN, vec1D_FLT* turnAngle, int* frame_len, int* frame_num,
                                          const std::string& dir_path, const int& polar_type, const int& data_type);

/* Loads the next sliding window of pulses onto the GPU (isar_imaging_export.h:73;
 * reference body isar_imaging_export.cpp:73-118). */
extern "C" DLL_EXPORT_API int dataExtracting(vec1D_INT* dataWFileSn, vec1D_DBL* dataNOut, vec1D_FLT* turnAngleOut,
                                             vec1D_COM_FLT* dataW, const vec1D_DBL& dataN, const vec1D_FLT& turnAngle,
                                             const int& frame_len, const int& frame_num, const int& sampling_stride,
                                             const int& window_head, const int& window_len, int& imaging_stride,
                                             const int& data_type);

/* Sizes and allocates everything once (isar_imaging_export.h:92; reference body
 * isar_imaging_export.cpp:121-205). */
extern "C" DLL_EXPORT_API void imagingMemInit(vec1D_FLT* img, vec1D_INT* dataWFileSn, vec1D_DBL* dataNOut,
                                              vec1D_FLT* turnAngleOut, vec1D_COM_FLT* dataW, const int& window_len,
                                              const int& frame_len, const int& data_type, const bool& if_hpc,
                                              const bool& if_hrrp);

/* One image of echo_num x 512 floats into h_img, synchronous (isar_imaging_export.h:108;
 * reference body isar_imaging_export.cpp:208-212 -> isar_main.cu:3-227). */
extern "C" DLL_EXPORT_API void isarMainSingle(float* h_img, const int& data_type, const int& option_alignment,
                                              const int& option_phase, const bool& if_hrrp, const bool& if_hpc,
                                              const bool& if_mtrc);

/* Text dump, one value per line with 5 decimals (isar_imaging_export.h:119; common.cu:1238-1252). */
extern "C" DLL_EXPORT_API void writeFileFLT(const std::string& outFilePath, const float* data, const size_t& data_size);

/* Releases everything imagingMemInit allocated (isar_imaging_export.h:130). */
extern "C" DLL_EXPORT_API void imagingMemDest(const int& data_type, const bool& if_hpc, const bool& if_hrrp);

#endif /* ISAR_IMAGING_EXPORT_H_ */
