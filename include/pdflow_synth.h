/* Seeded synthetic RGB-D pair generator (host C, no GPU).
 *
 * The reference ships no input data (its apps read i1/i2/z1/z2 PNGs that are
 * not in the tree, main_scene_flow_impair.cpp:62-65), so every workload of this
 * library is rendered from a seed:
 *   - 6-12 textured planar patches (one of them a background plane covering the
 *     whole view) at depths 0.5-4.5 m, textures are band-limited noise (sums of
 *     random plane waves in the patch's own metric coordinates);
 *   - pinhole camera fx = fy = 525*cols/640, principal point ((cols-1)/2, (rows-1)/2);
 *   - frame 2: every patch moves rigidly (|t| <= 5 cm, rotation <= 2 deg about its
 *     centre) and is re-rendered, then N(0, 2^2) intensity noise is added;
 *   - 5% of the depth pixels of each frame are set to 0 (holes);
 *   - intensity is uint8, depth uint16 = round(metres * 5000).
 * Outputs are column-major (v + u*rows), the reference's layout
 * (scene_flow_impair.cpp:230-232).  The solver input planes are
 * I = float(u8) and Z = float(u16)/5000 (scene_flow_impair.cpp:232, :242).
 */
#ifndef PDFLOW_SYNTH_H
#define PDFLOW_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Renders one pair.  Any output pointer may be NULL.
 * i1,i2: rows*cols uint8; z1,z2: rows*cols uint16;
 * gt: 3*rows*cols floats, ground-truth metric motion of the frame-1 point seen at
 * each pixel, in the solver's output axes (gt[0*P..] depth axis, gt[1*P..]
 * horizontal, gt[2*P..] vertical; pdflow_cudalib.cu:1157-1159), 0 where frame 1
 * has no depth.  Returns 0 on success. */
int pdf_synth_pair(uint64_t seed, unsigned rows, unsigned cols,
                   uint8_t *i1, uint16_t *z1, uint8_t *i2, uint16_t *z2, float *gt);

/* Same, converted to the solver's float planes: I = float(u8), Z = float(u16)/5000. */
int pdf_synth_pair_float(uint64_t seed, unsigned rows, unsigned cols,
                         float *I1, float *Z1, float *I2, float *Z2, float *gt);

/* n pairs with seeds seed0 .. seed0+n-1 rendered on `threads` host threads into consecutive
 * rows*cols slots of each output (pair i at offset i*rows*cols; gt at i*3*rows*cols, may be
 * NULL); same planes as pdf_synth_pair_float. */
int pdf_synth_pairs_float(uint64_t seed0, unsigned n, unsigned rows, unsigned cols, float *I1,
                          float *Z1, float *I2, float *Z2, float *gt, int threads);

/* The zero-motion variant: frame 2 equals frame 1 exactly (no motion, no noise,
 * same holes).  Used by the zero-flow property tests. */
int pdf_synth_static_float(uint64_t seed, unsigned rows, unsigned cols,
                           float *I, float *Z);

/* Frame t of a moving sequence of the same scene: every patch's motion applied t times
 * (t = 0 and t = 1 render the scene of pdf_synth_pair without/with its frame-2 motion; frames
 * t >= 1 carry intensity noise).  Used for the streaming mode, one new frame per solve. */
int pdf_synth_sequence_float(uint64_t seed, unsigned rows, unsigned cols, unsigned t, float *I,
                             float *Z);

/* FNV-1a 64-bit hash of a byte buffer (used to pin generator outputs). */
uint64_t pdf_synth_hash(const void *data, uint64_t nbytes);

#ifdef __cplusplus
}
#endif
#endif
