/* vp_io.h -- host-side scene and image I/O of the volumetric path tracer
 * (libvolprim_host.so, plain C++, no GPU needed).
 *
 * Reference interfaces each entry point replaces (kfernandez31/cuda-volprim):
 *   vp_io_load_ply   loadPrimitivesFromPLY (src/thesis/host/utils/io/ply.cpp:34-175):
 *                    14 float vertex properties x y z rot_0..3 (w x y z)
 *                    scale_0..2 (log) albedo_0..2 sigma_t_0; scale = exp,
 *                    quaternion normalised, mass = sigma_t_0 * sigma_multiplier,
 *                    albedo clamped to [0, 1] (NaN -> 0), rejects scale <= 0,
 *                    mass <= 0 and non-finite values.  albedo_override >= 0
 *                    replaces every albedo (ply.cpp:89-106).
 *   vp_io_save_ply   the inverse (binary little-endian), for generated scenes
 *   vp_io_load_hdr   loadHDR (src/thesis/host/utils/io/hdr.cpp:32-163): Radiance
 *                    RGBE, "-Y H +X W" only, new-style RLE or flat scanlines,
 *                    value = byte * 2^(e - 136), alpha = 1
 *   vp_io_save_hdr   the inverse (flat RGBE scanlines), for generated maps
 *   vp_io_save_exr   saveExr (src/thesis/host/utils/io/exr.cpp:35-105): single-part
 *                    scanline OpenEXR, channels B G R FLOAT, no compression,
 *                    rows written bottom-up (the image is flipped vertically)
 *   vp_io_load_exr   reader of the files vp_io_save_exr writes (tests, tools)
 *
 * Buffers the library returns are released with vp_io_free.  Errors return a
 * non-zero code; vp_io_last_error() describes the last one on the thread.
 */
#ifndef VP_IO_H
#define VP_IO_H

#include <stddef.h>
#include <stdint.h>

#include "vp_api.h"

#ifdef __cplusplus
extern "C" {
#endif

int vp_io_load_ply(const char* path, float sigma_multiplier, const float albedo_override[3], vp_primitive** out,
                   uint32_t* n);
/* rows as vp_primitive; log-scale and sigma_t_0 = mass / sigma_multiplier are written */
int vp_io_save_ply(const char* path, const vp_primitive* prims, uint32_t n, float sigma_multiplier);
int vp_io_load_hdr(const char* path, float** rgba, uint32_t* width, uint32_t* height);
int vp_io_save_hdr(const char* path, const float* rgba, uint32_t width, uint32_t height);
int vp_io_save_exr(const char* path, const float* rgba, uint32_t width, uint32_t height);
int vp_io_load_exr(const char* path, float** rgba, uint32_t* width, uint32_t* height);
void vp_io_free(void* p);
const char* vp_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VP_IO_H */
