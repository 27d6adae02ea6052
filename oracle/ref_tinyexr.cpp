// ref_tinyexr.cpp — test-only shim over the reference's own EXR loader
// (extern/tiny_exr.h with extern/miniz.c, as src/Core/Image.cpp:57-72 uses it
// through LoadEXR), so the tests can check the .exr reader of
// stratum2_amd/assets.py against it, and write test files with each
// compression.  Built by oracle/Makefile into oracle/_ref/ only when the
// reference tree is present.
#include <cstdlib>
#include <cstring>
#include <vector>

#define TINYEXR_IMPLEMENTATION
#include "tiny_exr.h"

extern "C" {

// LoadEXR into out (capacity cap floats, RGBA per pixel); returns 0 on success
// and the extent in w, h, or tinyexr's (negative) error code.
int ref_exr_load(const char* path, float* out, unsigned long long cap, int* w, int* h) {
    float* rgba = nullptr;
    const char* err = nullptr;
    const int ret = LoadEXR(&rgba, w, h, path, &err);
    if (ret != TINYEXR_SUCCESS) {
        if (err) FreeEXRErrorMessage(err);
        return ret;
    }
    const unsigned long long n = 4ull * (unsigned long long)(*w) * (unsigned long long)(*h);
    if (n <= cap) std::memcpy(out, rgba, n * sizeof(float));
    std::free(rgba);
    return n <= cap ? 0 : -100;
}

// Writes nch float planes (names given as one string of single-letter channel
// names, e.g. "ABGR", in file order) with the given compression
// (TINYEXR_COMPRESSIONTYPE_*) and pixel type (TINYEXR_PIXELTYPE_HALF or _FLOAT).
int ref_exr_save(const char* path, const float* planes, int w, int h, const char* names, int compression,
                 int pixel_type) {
    const int nch = (int)std::strlen(names);
    EXRHeader header;
    InitEXRHeader(&header);
    EXRImage image;
    InitEXRImage(&image);
    std::vector<const float*> ptrs(nch);
    for (int c = 0; c < nch; c++) ptrs[c] = planes + (size_t)c * w * h;
    image.num_channels = nch;
    image.images = (unsigned char**)ptrs.data();
    image.width = w;
    image.height = h;
    header.num_channels = nch;
    std::vector<EXRChannelInfo> ch(nch);
    std::vector<int> pt(nch, TINYEXR_PIXELTYPE_FLOAT), rpt(nch, pixel_type);
    for (int c = 0; c < nch; c++) {
        std::memset(&ch[c], 0, sizeof(EXRChannelInfo));
        ch[c].name[0] = names[c];
    }
    header.channels = ch.data();
    header.pixel_types = pt.data();
    header.requested_pixel_types = rpt.data();
    header.compression_type = compression;
    const char* err = nullptr;
    const int ret = SaveEXRImageToFile(&image, &header, path, &err);
    if (err) FreeEXRErrorMessage(err);
    return ret;
}

}  // extern "C"
