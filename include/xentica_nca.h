/*
 * xentica_nca.h -- neural-CA surrogate kernels (gfx950).
 *
 * No counterpart in the reference: Xentica has no learned component.
 * These kernels serve the surrogate that predicts a coarse observable of
 * an exact model 16 steps ahead (the exact engine stays the ground truth): a
 * coarse-graining encoder from an engine's uint8 main plane, and the
 * forward and backward passes of a 3x3 depthwise perception (16 channels,
 * fp32) followed by a 1x1 MLP 16 -> 64 -> 16 on bf16 MFMA with fp32
 * accumulation.
 *
 * Layouts: fine field uint8 [batch][16h][16w] (value = byte & value_mask);
 * x, y, dy fp32 [batch][h][w][16]; pred and block_mean fp32
 * [batch][h][w]; dw fp32 [16][9] (tap (dy+1)*3 + (dx+1)), bdw fp32 [16],
 * w1 bf16 [64][16], b1 fp32 [64], w2 bf16 [16][64], b2 fp32 [16];
 * grads fp32 [XEN_NCA_PARAMS] = dw | bdw | w1 | b1 | w2 | b2.  The coarse
 * grid is a torus per sample; batch * h * w must be a multiple of 16.
 * Input channel 4*sy + sx of coarse cell (y, x) is the mean of fine
 * sub-block rows 16y + 4sy .. +3, columns 16x + 4sx .. +3.
 *
 * Same conventions as xentica_hip.h: 0 on success, xen_last_error() on
 * failure; all pointers are device pointers; work is queued on `stream`.
 */
#ifndef XENTICA_NCA_H
#define XENTICA_NCA_H

#include <stdint.h>

#include "xentica_hip.h"

#define XEN_NCA_CHANNELS 16
#define XEN_NCA_HIDDEN 64
#define XEN_NCA_PARAMS 2288

#ifdef __cplusplus
extern "C" {
#endif

/* Coarse-grain a fine field into x (16 sub-block means per coarse cell);
 * block_mean (may be NULL) receives the mean of each 16x16 block. */
int xen_nca_encode(const uint8_t *fine, long long batch, int h, int w, uint32_t value_mask,
                   float *x, float *block_mean, xen_stream stream);

/* y = network(x); pred (may be NULL) = mean of y's 16 channels. */
int xen_nca_forward(const float *x, long long batch, int h, int w, const float *dw,
                    const float *bdw, const uint16_t *w1, const float *b1, const uint16_t *w2,
                    const float *b2, float *y, float *pred, xen_stream stream);

/* xen_nca_encode followed by xen_nca_forward in one launch, without x:
 * y and pred (either may be NULL) equal theirs bit for bit.  h must be a
 * multiple of 8 and w of 16. */
int xen_nca_predict(const uint8_t *fine, long long batch, int h, int w, uint32_t value_mask,
                    const float *dw, const float *bdw, const uint16_t *w1, const float *b1,
                    const uint16_t *w2, const float *b2, float *y, float *pred,
                    xen_stream stream);

/* Rows of the fp32 workspace xen_nca_backward needs (XEN_NCA_PARAMS
 * floats each), or -1 for an invalid grid. */
long long xen_nca_partial_rows(long long batch, int h, int w);

/* grads = d(sum(dy * y)) / d(parameters), recomputing the forward pass;
 * partials is a caller-owned workspace of partial_rows rows. */
int xen_nca_backward(const float *x, const float *dy, long long batch, int h, int w,
                     const float *dw, const float *bdw, const uint16_t *w1, const float *b1,
                     const uint16_t *w2, const float *b2, float *partials,
                     long long partial_rows, float *grads, xen_stream stream);

/* xen_nca_backward without the caller's workspace and without allocating
 * anything: the partial rows live in a static device array of the library
 * (256 rows of XEN_NCA_PARAMS floats, 2.3 MB per GPU, held until the
 * process ends; a grid that would need more rows is rejected, which no
 * grid does today since the backward launch caps its workgroups at 256).
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  Every call on one GPU shares that GPU's
 * workspace: calls that overlap on one GPU (different streams or host
 * threads) race and must be serialised by the caller.  Results equal
 * xen_nca_backward's bit for bit. */
int xen_nca_backward_noalloc(const float *x, const float *dy, long long batch, int h, int w,
                             const float *dw, const float *bdw, const uint16_t *w1,
                             const float *b1, const uint16_t *w2, const float *b2,
                             float *grads, xen_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* XENTICA_NCA_H */
