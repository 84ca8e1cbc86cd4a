/* Test infrastructure only: the reference's PNG read step, xSeqPNG::xBackendRead
   (libs/pmbb-spng/src/xSeqPNG.cpp:91-148), as a C function over the reference's own libspng
   (thirdparty/spng/src/spng.c, compiled where it lies by oracle/Makefile). It decodes one file to
   RGB8 with decode flags 0 after the size and bit-depth checks, and stores the planes R, G, B. The
   tests compare the engine's PNG decoder with it. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spng.h"

/* 0 on success; otherwise a non-zero code and a message in err. */
int spng_ref_decode(const char* path, int32_t width, int32_t height, uint8_t* planar, char* err, int errlen) {
  FILE* f = fopen(path, "rb");
  if (!f) {
    snprintf(err, (size_t)errlen, "cannot open");
    return 1;
  }
  spng_ctx* ctx = spng_ctx_new(0);
  int rc = spng_set_png_file(ctx, f);
  size_t size_rgba = 0;
  struct spng_ihdr ihdr;
  if (!rc) rc = spng_decode_chunks(ctx);
  if (!rc) rc = spng_decoded_image_size(ctx, SPNG_FMT_RGBA8, &size_rgba);
  if (!rc) rc = spng_get_ihdr(ctx, &ihdr);
  if (rc) {
    snprintf(err, (size_t)errlen, "%s", spng_strerror(rc));
    spng_ctx_free(ctx);
    fclose(f);
    return 2;
  }
  const char* mismatch = (int32_t)ihdr.width != width ? "Width does not match"
                         : (int32_t)ihdr.height != height ? "Height does not match"
                         : ihdr.bit_depth != 8 ? "BitDepth does not match" : NULL;
  if (mismatch) {
    snprintf(err, (size_t)errlen, "%s", mismatch);
    spng_ctx_free(ctx);
    fclose(f);
    return 3;
  }
  uint8_t* rgb = (uint8_t*)malloc(size_rgba);
  rc = spng_decode_image(ctx, rgb, size_rgba, SPNG_FMT_RGB8, 0);
  spng_ctx_free(ctx);
  fclose(f);
  if (rc) {
    snprintf(err, (size_t)errlen, "%s", spng_strerror(rc));
    free(rgb);
    return 4;
  }
  const size_t n = (size_t)width * (size_t)height;
  for (size_t i = 0; i < n; i++) {
    planar[i] = rgb[3 * i];
    planar[n + i] = rgb[3 * i + 1];
    planar[2 * n + i] = rgb[3 * i + 2];
  }
  free(rgb);
  return 0;
}
