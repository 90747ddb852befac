/* PNG input/output for the command-line pair evaluator (host C, zlib; part of
 * libpdflow_png.so and of the pdflow-impair executable).
 *
 * The reference app reads its four frames with OpenCV (scene_flow_impair.cpp:218-277: intensity
 * with CV_LOAD_IMAGE_GRAYSCALE, depth unchanged then convertTo(CV_32FC1, 1/5000)) and writes the
 * RGB representation with cv::imwrite (:322-342).  OpenCV is not part of this build; this codec
 * covers the formats the reference's README asks for (8-bit grey intensity, 16-bit grey depth
 * scaled by 5000; README.txt:59-66), plus 8/16-bit grey, grey+alpha, RGB, RGBA and 8-bit palette
 * images without interlacing.  Colour images read as grey use the ITU-R BT.601 weights.
 */
#ifndef PDFLOW_PNG_H
#define PDFLOW_PNG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pdf_png_status {
    PDF_PNG_OK = 0,
    PDF_PNG_ERR_IO = 1,          /* cannot open / read / write the file */
    PDF_PNG_ERR_FORMAT = 2,      /* not a PNG, bad chunk, bad CRC, bad zlib stream */
    PDF_PNG_ERR_UNSUPPORTED = 3, /* interlaced, or a colour type / bit depth outside the list above */
    PDF_PNG_ERR_NOMEM = 4
};

/* Decodes a PNG into interleaved samples, row-major: channels = 1 (grey), 2 (grey+alpha), 3 (RGB)
 * or 4 (RGBA; palette images come back as RGB or RGBA), bit_depth 8 or 16 (16-bit samples in host
 * order).  *data is malloc'd; release with pdf_png_free. */
int pdf_png_read(const char *path, unsigned *width, unsigned *height, unsigned *channels,
                 unsigned *bit_depth, void **data);
void pdf_png_free(void *data);

/* Encodes row-major interleaved samples (channels 1 or 3, bit_depth 8 or 16, 16-bit in host
 * order) as a non-interlaced PNG. */
int pdf_png_write(const char *path, unsigned width, unsigned height, unsigned channels,
                  unsigned bit_depth, const void *data);

/* The reference's frame loading (scene_flow_impair.cpp:218-277) into column-major solver planes
 * of width*height floats: I = intensity as 8-bit grey (16-bit images keep their high byte, colour
 * images are converted with BT.601), Z = depth sample * (1/5000) (8- or 16-bit grey).  Either
 * path may be NULL.  The image sizes must be equal to width x height. */
int pdf_load_frame(const char *intensity_path, const char *depth_path, unsigned width,
                   unsigned height, float *I, float *Z);

/* Size of a PNG without decoding it. */
int pdf_png_size(const char *path, unsigned *width, unsigned *height);

#ifdef __cplusplus
}
#endif
#endif
