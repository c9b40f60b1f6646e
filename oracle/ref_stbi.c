/* ref_stbi.c — TEST INFRASTRUCTURE.  Compiles the reference's own image
 * decoder, libs/stb_image/stb_image.h (the stbi_load call of
 * src/RayTracer/LumenScene.cpp:207), from where it lies in the reference
 * tree into oracle/_ref/libref_stbi.so, so that the tests can check
 * lt_image_load_rgba8 against it.  Built only where the reference tree is
 * present; nothing in the product loads it. */
#define STB_IMAGE_IMPLEMENTATION
#define STBI_ONLY_JPEG
#include <stddef.h>
#include <string.h>

#include "stb_image.h"

/* 0 on success; out == NULL returns the size only */
int ref_stbi_load_rgba8(const char* path, unsigned char* out, size_t cap, int* w, int* h) {
    int n = 0;
    unsigned char* d = stbi_load(path, w, h, &n, 4);
    if (!d) return -1;
    const size_t need = (size_t)(*w) * (size_t)(*h) * 4u;
    int rc = 0;
    if (out) {
        if (cap < need)
            rc = -2;
        else
            memcpy(out, d, need);
    }
    stbi_image_free(d);
    return rc;
}
