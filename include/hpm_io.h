/*
 * hpm_io.h — the reference's on-disk formats, without OpenCV.
 *
 *   cams/%08d_cam.txt  ReadCamera (HPM-MVS/HPM.cpp:160-185): "extrinsic",
 *                      4x4 [R|t], "intrinsic", 3x3 K, then
 *                      depth_min interval depth_num depth_max.
 *   pair.txt           GenerateSampleList (HPM-MVS/main.cpp:4-33): N, then
 *                      per view: ref_id, n, n x (id score); score <= 0 skipped.
 *   *.dmb              readDepthDmb/writeDepthDmb/readNormalDmb/writeNormalDmb
 *                      (HPM-MVS/HPM.cpp:270-401): int32 type=1, h, w, nb, then
 *                      h*w*nb float32 row-major.
 *   images             8-bit grey, PGM (P5) here; the reference decodes JPEG
 *                      with cv::imread(IMREAD_GRAYSCALE) (HPM.cpp:492).
 * All functions return 0 on success and a negative value on failure; none
 * exits the process.
 */
#ifndef HPM_IO_H_
#define HPM_IO_H_

#include <stddef.h>
#include <stdint.h>

#include "hpm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ReadCamera (HPM.cpp:160-185); width/height are left 0. */
int hpm_read_camera(const char* path, hpm_camera* cam);
/* Writer of the same layout (interval 0, depth_num 192). */
int hpm_write_camera(const char* path, const hpm_camera* cam);

/* GenerateSampleList (main.cpp:4-33). Call with ref_ids == NULL to get the
 * problem count. src_ids is max_src entries per problem; sources whose
 * score is <= 0 are dropped as in the reference. */
int hpm_read_pair(const char* path, int* num_problems, int* ref_ids, int* num_src, int* src_ids, int max_src);

/* read/write{Depth,Normal}Dmb (HPM.cpp:270-401). Call with data == NULL to
 * read only the header. type != 1 is an error (HPM.cpp:315-318). */
int hpm_read_dmb(const char* path, float* data, int* width, int* height, int* channels);
int hpm_write_dmb(const char* path, const float* data, int width, int height, int channels);

/* 8-bit binary PGM (P5). hpm_read_pgm_f32 converts to float (the
 * convertTo(CV_32FC1) of HPM.cpp:494); data == NULL reads only the size. */
int hpm_read_pgm_f32(const char* path, float* data, int* width, int* height);
int hpm_write_pgm_u8(const char* path, const uint8_t* data, int width, int height);

/* cv::imread(path, cv::IMREAD_GRAYSCALE) + convertTo(CV_32FC1)
 * (HPM.cpp:492-494) for JPEG and PGM files, chosen by content: for a JPEG
 * the luminance component as libjpeg's grayscale output computes it
 * (integer inverse DCT); baseline / extended sequential Huffman, 8-bit, grey
 * or YCbCr only. data == NULL reads only the size. Returns 0 or -1. */
int hpm_read_image_grey(const char* path, float* data, int* width, int* height);
/* The same decode of a JPEG in memory into width * height bytes; out ==
 * NULL returns the size. */
int hpm_decode_jpeg_grey(const uint8_t* data, size_t size, uint8_t* out, int* width, int* height);

/* cv::imread(path, cv::IMREAD_COLOR) (main.cpp:523, RunFusion's colours)
 * for JPEG and PGM files: width * height RGB byte triples (the reference's
 * BGR Mat in RGB order). JPEG chroma subsampled 2x1 or 2x2 is upsampled
 * with libjpeg's default ("fancy") triangle filter, YCbCr converted in its
 * fixed point; grey JPEGs and PGMs give the grey value in all channels.
 * data == NULL reads only the size. Returns 0 or -1. */
int hpm_read_image_rgb(const char* path, uint8_t* data, int* width, int* height);
/* The same decode of a JPEG in memory into width * height * 3 bytes; out ==
 * NULL returns the size. */
int hpm_decode_jpeg_rgb(const uint8_t* data, size_t size, uint8_t* out, int* width, int* height);

/* cv::resize(..., INTER_LINEAR) of a float image: half-pixel centres,
 * src = (dst + 0.5) * (sw / dw) - 0.5, edge-clamped, exact float weights.
 * Used for the >max_image_size downscale (HPM.cpp:524-551) and the prior
 * hierarchy (main.cpp:84-86, 238-240). channels interleaved. */
void hpm_resize_bilinear_c(const float* src, int sw, int sh, int channels, float* dst, int dw, int dh);
void hpm_resize_bilinear(const float* src, int sw, int sh, float* dst, int dw, int dh);

#ifdef __cplusplus
}
#endif

#endif /* HPM_IO_H_ */
