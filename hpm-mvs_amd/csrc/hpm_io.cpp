/*
 * hpm_io.cpp — readers and writers of the reference's folder layout
 * (include/hpm_io.h) and the bilinear resize used by the input downscale and
 * the prior hierarchy.
 */
#include "hpm_io.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <string>
#include <vector>

extern "C" int hpm_read_camera(const char* path, hpm_camera* cam)
{
    if (!path || !cam) {
        return -1;
    }
    std::ifstream file(path);
    if (!file) {
        return -2;
    }
    memset(cam, 0, sizeof(*cam));
    std::string word;
    file >> word; /* "extrinsic" */
    for (int i = 0; i < 3; ++i) {
        file >> cam->R[3 * i + 0] >> cam->R[3 * i + 1] >> cam->R[3 * i + 2] >> cam->t[i];
    }
    float tmp[4];
    file >> tmp[0] >> tmp[1] >> tmp[2] >> tmp[3];
    file >> word; /* "intrinsic" */
    for (int i = 0; i < 3; ++i) {
        file >> cam->K[3 * i + 0] >> cam->K[3 * i + 1] >> cam->K[3 * i + 2];
    }
    float interval = 0.0f, depth_num = 0.0f;
    file >> cam->depth_min >> interval >> depth_num >> cam->depth_max;
    if (file.fail()) {
        return -2;
    }
    return 0;
}

extern "C" int hpm_write_camera(const char* path, const hpm_camera* cam)
{
    if (!path || !cam) {
        return -1;
    }
    FILE* f = fopen(path, "w");
    if (!f) {
        return -2;
    }
    fprintf(f, "extrinsic\n");
    for (int i = 0; i < 3; ++i) {
        fprintf(f, "%.9g %.9g %.9g %.9g\n", cam->R[3 * i], cam->R[3 * i + 1], cam->R[3 * i + 2], cam->t[i]);
    }
    fprintf(f, "0 0 0 1\n\nintrinsic\n");
    for (int i = 0; i < 3; ++i) {
        fprintf(f, "%.9g %.9g %.9g\n", cam->K[3 * i], cam->K[3 * i + 1], cam->K[3 * i + 2]);
    }
    fprintf(f, "\n%.9g %.9g %d %.9g\n", cam->depth_min, 0.0, 192, cam->depth_max);
    return fclose(f) == 0 ? 0 : -2;
}

extern "C" int hpm_read_pair(const char* path, int* num_problems, int* ref_ids, int* num_src, int* src_ids,
                             int max_src)
{
    if (!path || !num_problems) {
        return -1;
    }
    std::ifstream file(path);
    if (!file) {
        return -2;
    }
    int n = 0;
    file >> n;
    if (file.fail() || n < 0) {
        return -2;
    }
    if (!ref_ids) {
        *num_problems = n;
        return 0;
    }
    if (*num_problems < n || !num_src || !src_ids || max_src <= 0) {
        return -1;
    }
    for (int i = 0; i < n; ++i) {
        int ref = 0, ns = 0;
        file >> ref >> ns;
        if (file.fail() || ns < 0) {
            return -2;
        }
        int kept = 0;
        for (int j = 0; j < ns; ++j) {
            int id = 0;
            float score = 0.0f;
            file >> id >> score;
            if (file.fail()) {
                return -2;
            }
            if (score <= 0.0f) {
                continue;
            }
            if (kept < max_src) {
                src_ids[(size_t)i * max_src + kept] = id;
            }
            ++kept;
        }
        ref_ids[i] = ref;
        num_src[i] = kept < max_src ? kept : max_src;
    }
    *num_problems = n;
    return 0;
}

extern "C" int hpm_read_dmb(const char* path, float* data, int* width, int* height, int* channels)
{
    if (!path) {
        return -1;
    }
    FILE* f = fopen(path, "rb");
    if (!f) {
        return -2;
    }
    int32_t hdr[4] = {-1, 0, 0, 0};
    if (fread(hdr, sizeof(int32_t), 4, f) != 4 || hdr[0] != 1 || hdr[1] < 0 || hdr[2] < 0 || hdr[3] < 1) {
        fclose(f);
        return -2;
    }
    if (width) {
        *width = hdr[2];
    }
    if (height) {
        *height = hdr[1];
    }
    if (channels) {
        *channels = hdr[3];
    }
    if (data) {
        const size_t n = (size_t)hdr[1] * hdr[2] * hdr[3];
        if (fread(data, sizeof(float), n, f) != n) {
            fclose(f);
            return -2;
        }
    }
    fclose(f);
    return 0;
}

extern "C" int hpm_write_dmb(const char* path, const float* data, int width, int height, int channels)
{
    if (!path || !data || width < 0 || height < 0 || channels < 1) {
        return -1;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        return -2;
    }
    const int32_t hdr[4] = {1, height, width, channels};
    const size_t n = (size_t)width * height * channels;
    const bool ok = fwrite(hdr, sizeof(int32_t), 4, f) == 4 && fwrite(data, sizeof(float), n, f) == n;
    return (fclose(f) == 0 && ok) ? 0 : -2;
}

static int pgm_token(FILE* f, int* out)
{
    int c = fgetc(f);
    while (c != EOF) {
        if (c == '#') {
            while (c != EOF && c != '\n') {
                c = fgetc(f);
            }
        } else if (!isspace(c)) {
            break;
        }
        c = fgetc(f);
    }
    if (c == EOF || !isdigit(c)) {
        return -1;
    }
    int v = 0;
    while (c != EOF && isdigit(c)) {
        v = v * 10 + (c - '0');
        c = fgetc(f);
    }
    *out = v; /* the single whitespace after the token is consumed */
    return 0;
}

extern "C" int hpm_read_pgm_f32(const char* path, float* data, int* width, int* height)
{
    if (!path) {
        return -1;
    }
    FILE* f = fopen(path, "rb");
    if (!f) {
        return -2;
    }
    char magic[2];
    int w = 0, h = 0, maxv = 0;
    if (fread(magic, 1, 2, f) != 2 || magic[0] != 'P' || magic[1] != '5' || pgm_token(f, &w) != 0 ||
        pgm_token(f, &h) != 0 || pgm_token(f, &maxv) != 0 || w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) {
        fclose(f);
        return -2;
    }
    if (width) {
        *width = w;
    }
    if (height) {
        *height = h;
    }
    if (data) {
        std::vector<uint8_t> buf((size_t)w * h);
        if (fread(buf.data(), 1, buf.size(), f) != buf.size()) {
            fclose(f);
            return -2;
        }
        for (size_t i = 0; i < buf.size(); ++i) {
            data[i] = (float)buf[i];
        }
    }
    fclose(f);
    return 0;
}

extern "C" int hpm_write_pgm_u8(const char* path, const uint8_t* data, int width, int height)
{
    if (!path || !data || width <= 0 || height <= 0) {
        return -1;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        return -2;
    }
    fprintf(f, "P5\n%d %d\n255\n", width, height);
    const size_t n = (size_t)width * height;
    const bool ok = fwrite(data, 1, n, f) == n;
    return (fclose(f) == 0 && ok) ? 0 : -2;
}

extern "C" void hpm_resize_bilinear_c(const float* src, int sw, int sh, int channels, float* dst, int dw, int dh)
{
    const double sx_scale = (double)sw / dw;
    const double sy_scale = (double)sh / dh;
    for (int y = 0; y < dh; ++y) {
        float fy = (float)((y + 0.5) * sy_scale - 0.5);
        int y0 = (int)floorf(fy);
        fy -= (float)y0;
        int y1 = y0 + 1;
        if (y0 < 0) {
            y0 = 0;
            y1 = 0;
            fy = 0.0f;
        }
        if (y1 >= sh) {
            y0 = sh - 1;
            y1 = sh - 1;
            fy = 0.0f;
        }
        for (int x = 0; x < dw; ++x) {
            float fx = (float)((x + 0.5) * sx_scale - 0.5);
            int x0 = (int)floorf(fx);
            fx -= (float)x0;
            int x1 = x0 + 1;
            if (x0 < 0) {
                x0 = 0;
                x1 = 0;
                fx = 0.0f;
            }
            if (x1 >= sw) {
                x0 = sw - 1;
                x1 = sw - 1;
                fx = 0.0f;
            }
            for (int c = 0; c < channels; ++c) {
                const float a = src[((size_t)y0 * sw + x0) * channels + c];
                const float b = src[((size_t)y0 * sw + x1) * channels + c];
                const float d = src[((size_t)y1 * sw + x0) * channels + c];
                const float e = src[((size_t)y1 * sw + x1) * channels + c];
                const float top = a * (1.0f - fx) + b * fx;
                const float bot = d * (1.0f - fx) + e * fx;
                dst[((size_t)y * dw + x) * channels + c] = top * (1.0f - fy) + bot * fy;
            }
        }
    }
}

extern "C" void hpm_resize_bilinear(const float* src, int sw, int sh, float* dst, int dw, int dh)
{
    hpm_resize_bilinear_c(src, sw, sh, 1, dst, dw, dh);
}
