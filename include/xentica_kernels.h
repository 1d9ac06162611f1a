/*
 * xentica_kernels.h -- hand-written gfx950 kernels of the fast engines.
 *
 * These entry points have no one-to-one counterpart in the reference: it
 * runs every model through the generated emit/absorb kernels launched
 * from CellularAutomaton.step() (xentica/core/base.py:703-728), with the
 * state hashed on the host by binascii.crc32 in its tests
 * (tests/core/test_base.py:39-41).  Each function below computes the same
 * results as that path for the models it accepts; the comment of each
 * names the reference behaviour it reproduces.
 *
 * Same conventions as xentica_hip.h: 0 on success, xen_last_error() on
 * failure; all pointers are device pointers unless stated otherwise.
 */
#ifndef XENTICA_KERNELS_H
#define XENTICA_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#include "xentica_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- bit-packed Life-like engine (2-D Moore torus) ----
 * Replaces the emit+absorb pair of GameOfLife / LifelikeCA
 * (examples/game_of_life.py:42-70, 190-197) for 1-bit states:
 * next = ((born >> count) & 1) | ((sustain >> count) & 1) & state.
 * State: 1 bit per cell, 32 cells per uint32 word, x0 fastest;
 * width_words = width / 32.  One launch advances `depth` (1..8) steps;
 * the rows are split into `strips` balanced strips, one wavefront per
 * (62-word column tile, strip); the words left over after the whole
 * tiles run in row segments of ceil(height / strips) rows, up to four
 * interior segments sharing a wavefront.  strips = 0 picks one, two or
 * three waves per SIMD (at most one strip per row) by the library's cost
 * model.  The field must be smaller than 2 GiB. */
int xen_life_bits_launch(const void *src, void *dst, int width_words, int height,
                         uint32_t born, uint32_t sustain, int depth, int strips,
                         xen_stream stream);
/* Advance `steps` steps ping-ponging between buf_a (holding the state) and
 * buf_b; *result_in_b tells which buffer holds the result. */
int xen_life_bits_run(void *buf_a, void *buf_b, int width_words, int height,
                      uint32_t born, uint32_t sustain, long long steps, int depth,
                      int strips, int *result_in_b, xen_stream stream);

/* ---- bit-sliced batch of Life-like tori of any width ----
 * Replaces one emit+absorb launch pair per instance per step of the
 * reference's LifelikeCA (examples/game_of_life.py:183-203) for many
 * independent 2-D Moore tori at once: instance i (width w_i >= 32, any
 * height, its own birth/survival masks) is stored at word base_i as rows of
 * pitch_i = ceil(w_i / 32) uint32 words (cell x of a row in bit x % 32 of
 * word x / 32, the bits past w_i zero).  One launch advances every instance
 * `depth` (1..8) steps from src to dst (the whole batch, total_bytes < 2
 * GiB).  insts: two int4 per instance, (base_i, pitch_i, w_i, h_i) and
 * (born_i, sustain_i, 0, 0).  lanes: 64 int4 per wave, one per lane:
 * (instance or -1 for an idle lane, extended word column c -- the 32 cells
 * from cell 32 c modulo w --, first output row, output rows or 0 for a lane
 * that stores nothing); wave_rows[w]: the most output rows of any lane of
 * wave w.  The lanes of a column tile of at most 62 stored words sit next
 * to each other with one halo lane on each side (c = first - 1, last + 1). */
int xen_life_batch_launch(const void *src, void *dst, long long total_bytes, const void *insts,
                          const void *lanes, const int *wave_rows, int n_waves, int depth,
                          xen_stream stream);
/* Row-padded packing of one instance: a uint8 plane of width x height (0/1
 * values) into rows of pitch = ceil(width / 32) words, and back into the
 * plane and/or int3 colours by state (color_dead / color_alive: host
 * pointers to three ints; cells or colors may be NULL). */
int xen_bits_pack_rows(const uint8_t *cells, uint32_t *bits, int width, int height, int pitch,
                       xen_stream stream);
int xen_bits_expand_rows(const uint32_t *bits, uint8_t *cells, int32_t *colors, int width,
                         int height, int pitch, const int32_t *color_dead,
                         const int32_t *color_alive, xen_stream stream);

/* ---- 3-D Life-like engines (torus, the reference's 3-D Moore list) ----
 * The reference's 3-D MooreNeighborhood keeps the centre and drops (1,1,1)
 * (xentica/core/topology/neighborhood.py:159-167); the count of a cell is
 * its 3x3x3 cube minus the cell at (x+1, y+1, z+1) and
 * next = ((born >> count) & 1) | ((sustain >> count) & 1) & state.
 * Bit-packed: 32 cells per uint32, x fastest, width / 32 dividing 64,
 * field < 2 GiB; z_tiles = 0 picks about four waves per SIMD.  One call
 * advances one step from src to dst. */
int xen_life3d_bits_step(const void *src, void *dst, int width, int height, int depth,
                         uint32_t born, uint32_t sustain, int z_tiles, xen_stream stream);
/* Same rule on one uint8 per cell, one thread per cell reading its 26
 * neighbours directly: an independent restatement for cross-checks, any
 * width. */
int xen_life3d_u8_direct_step(const void *src, void *dst, int width, int height, int depth,
                              uint32_t born, uint32_t sustain, xen_stream stream);

/* ---- layout conversion between the reference plane and packed bits ----
 * cells: the reference's main-state plane (one uint8 per cell, value 0/1,
 * xentica/core/base.py:659-661); bits: 32 cells per uint32, x0 fastest.
 * n_cells must be a multiple of 32. */
int xen_bits_pack(const uint8_t *cells, uint32_t *bits, long long n_cells, xen_stream stream);
int xen_bits_unpack(const uint32_t *bits, uint8_t *cells, long long n_cells, xen_stream stream);

/* A packed plane expanded into the reference's arrays in one pass: cells
 * (may be NULL) receives one 0/1 byte per cell, colors (may be NULL) the
 * int3 colour of each cell, color_dead[0..2] for dead and color_alive[0..2]
 * for live cells (host pointers).  Replaces reading the main plane and the
 * colours that absorb() wrote (xentica/core/base.py:659-661,
 * xentica/core/color_effects.py:66-82) for a model whose colour is a
 * function of its 1-bit state.  n_cells a multiple of 32; cells 4-byte and
 * colors 16-byte aligned. */
int xen_bits_expand(const uint32_t *bits, uint8_t *cells, int32_t *colors, long long n_cells,
                    const int32_t *color_dead, const int32_t *color_alive, xen_stream stream);

/* The low byte of each of n int32 values into dst: the
 * `img.get().astype(numpy.uint8)` of the reference's render()
 * (xentica/core/base.py:753) done on the device, so that a quarter of the
 * bytes cross to the host.  src 16-byte aligned, dst 4-byte aligned. */
int xen_low_bytes(const int32_t *src, uint8_t *dst, long long n, xen_stream stream);

/* ---- device CRC-32 (the binascii.crc32 of tests/core/test_base.py:39-41) ----
 * Zero-initialised, un-inverted CRC-32 register of every `chunk`-byte
 * chunk of a byte plane (the last chunk may be shorter) into
 * partials[ceil(n / chunk)]; the host combines them into crc32 of the
 * whole plane.  The _bits variant hashes the uint8 expansion (one byte
 * 0/1 per cell) of a packed plane without materialising it; its chunk
 * must be a multiple of 32. */
int xen_crc32_u8_partials(const uint8_t *data, long long n, long long chunk,
                          uint32_t *partials, xen_stream stream);
int xen_crc32_bits_partials(const uint32_t *bits, long long n_cells, long long chunk,
                            uint32_t *partials, xen_stream stream);

/* ---- sum of one field over a plane of container words ----
 * No reference counterpart: the reference checks conserved quantities on
 * the host after `cells.get()` (examples/conserved_life.py:237-256).
 * partials[b] = sum over the cells block b visits (grid-stride, n_partials
 * blocks of 256 threads) of (word >> shift) & (2^width - 1); the host adds
 * the n_partials values.  word_bytes is 1, 2, 4, 8 or 16 (two little-endian
 * 64-bit halves), width 1..32, shift + width within the word; words aligned
 * to their size (16-byte words to 16). */
int xen_field_sum_partials(const void *words, long long n, int word_bytes, int shift, int width,
                           uint64_t *partials, int n_partials, xen_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* XENTICA_KERNELS_H */
