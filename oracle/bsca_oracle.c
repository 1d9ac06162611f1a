/*
 * bsca_oracle.c -- CPU restatement of the reference's buffered-state step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU engines are
 * checked against and the CPU baseline of the benchmark; no product code
 * links or calls it.
 *
 * It follows the reference's plain structure literally
 * (xentica/core/base.py:254-283, 703-728): a flat array fld of
 * n*(k+1)+1 words, plane 0 the main state and plane p buffer slot p; an
 * emit loop over all cells writing every buffer slot, then an absorb loop
 * over all cells reading neighbour slots and writing the main state and
 * the colour.  Neighbour j of cell x is x + delta[j]; buffers[q] goes to
 * slot rev[q]+1 and neighbors[j].buffer comes from slot j+1
 * (xentica/core/base.py:223-236).  Torus: if any coordinate of a
 * neighbour is off the board, every coordinate becomes (c + w) % w
 * (xentica/core/topology/border.py:104-117); static border: the whole
 * off-board word is the constant (border.py:144-151,
 * xentica/core/properties.py:467-483).  Cells are indexed x0 fastest
 * (xentica/core/topology/lattice.py:146-204).
 *
 * Models restated (each cites the reference model it follows):
 *   oracle_lifelike      GameOfLife / LifelikeCA / GameOfLife6D / 3-D
 *                        Life-like / Life-like batches
 *                        (examples/game_of_life.py:15-90, 183-203)
 *   oracle_color_life    GameOfLifeColor (examples/game_of_life.py:112-159)
 *   oracle_shifting_sands ShiftingSands (examples/shifting_sands.py:15-47)
 *   oracle_noisetv       NoiseTV (examples/noisetv.py:10-28)
 *   oracle_energy_life   EnergyLife (no reference counterpart: the
 *                        project's EvoLife-style model on a 64-bit word,
 *                        and its 128-bit variant, written in the
 *                        reference's DSL; the rule is restated here from
 *                        its definition)
 *   oracle_xorshift32_period  period of the xorshift32 map from a state
 *   oracle_crc32         CRC-32 of binascii.crc32 (tests/core/test_base.py:39-41)
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_DIMS 16

typedef struct {
    int ndim;
    int64_t size[MAX_DIMS];
    int64_t n;
    int k;
    const int8_t *delta;   /* k * ndim offsets */
    const int32_t *rev;    /* k entries */
    int border_static;     /* 0 = torus, 1 = static */
    int64_t border_value;  /* packed word of an off-board cell */
} lattice_t;

static int setup(lattice_t *lat, int ndim, const int64_t *sizes, int k,
                 const int8_t *delta, const int32_t *rev, int border_static,
                 int64_t border_value) {
    if (ndim < 1 || ndim > MAX_DIMS) return 1;
    lat->ndim = ndim;
    lat->n = 1;
    for (int a = 0; a < ndim; ++a) {
        lat->size[a] = sizes[a];
        lat->n *= sizes[a];
    }
    lat->k = k;
    lat->delta = delta;
    lat->rev = rev;
    lat->border_static = border_static;
    lat->border_value = border_value;
    return 0;
}

/* Flat index of neighbour j of cell i, or -1 when it is off the board
 * under a static border (lattice.py:146-221, border.py:104-117). */
static int64_t neighbor_index(const lattice_t *lat, int64_t i, int j) {
    int64_t x[MAX_DIMS];
    int64_t rest = i;
    for (int a = 0; a < lat->ndim; ++a) {
        if (a < lat->ndim - 1) {
            x[a] = rest % lat->size[a];
            rest /= lat->size[a];
        } else {
            x[a] = rest;
        }
    }
    int off = 0;
    for (int a = 0; a < lat->ndim; ++a) {
        x[a] += lat->delta[j * lat->ndim + a];
        if (x[a] < 0 || x[a] >= lat->size[a]) off = 1;
    }
    if (off) {
        if (lat->border_static) return -1;
        for (int a = 0; a < lat->ndim; ++a)
            x[a] = (x[a] + lat->size[a]) % lat->size[a];
    }
    int64_t idx = 0, stride = 1;
    for (int a = 0; a < lat->ndim; ++a) {
        idx += x[a] * stride;
        stride *= lat->size[a];
    }
    return idx;
}

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* MovingAverage (xentica/core/color_effects.py:106-130): scale by
 * SMOOTH_FACTOR, clamp to [old - FADE_OUT, old + FADE_IN]. */
static inline int32_t fade(int32_t v, int32_t old, int fade_in, int fade_out,
                           int smooth) {
    v *= smooth;
    int32_t hi = old + fade_in, lo = old - fade_out;
    if (v > hi) v = hi;
    if (v < lo) v = lo;
    return v;
}

/*
 * Life-like step on a 1-bit state stored in uint8 words.  The absorb
 * rule is the reference's: born = (B >> count) & 1,
 * sustained = (S >> count) & 1, state = born | sustained & state
 * (GameOfLife: B=8, S=12; LifelikeCA: B=rule, S=rule>>9).
 * col may be NULL (colours not kept); otherwise colour = state*255 per
 * channel through MovingAverage.
 * hashes, if not NULL, receives crc32 of plane 0 after every step.
 */
uint32_t oracle_crc32(uint32_t crc, const uint8_t *buf, int64_t len);

int oracle_lifelike(uint8_t *fld, int32_t *col, int ndim, const int64_t *sizes,
                    int k, const int8_t *delta, const int32_t *rev,
                    int border_static, int64_t border_value,
                    uint32_t born, uint32_t sustain,
                    int fade_in, int fade_out, int smooth,
                    int steps, int threads, uint32_t *hashes) {
    lattice_t lat;
    if (setup(&lat, ndim, sizes, k, delta, rev, border_static, border_value)) return 1;
    set_threads(threads);
    const int64_t n = lat.n;
    const uint8_t off_state = (uint8_t)border_value;
    for (int s = 0; s < steps; ++s) {
        /* emit: every buffer slot gets the (masked) state */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint8_t st = fld[i] & 1;
            for (int q = 0; q < k; ++q) fld[i + n * (lat.rev[q] + 1)] = st & 1;
        }
        /* absorb: count neighbours' slots, apply the rule, colour */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t count = 0;
            for (int j = 0; j < k; ++j) {
                int64_t idx = neighbor_index(&lat, i, j);
                uint8_t w = idx < 0 ? off_state : fld[idx + n * (j + 1)];
                count += w & 1;
            }
            uint32_t st = fld[i] & 1;
            uint32_t is_born = (born >> count) & 1;
            uint32_t is_sus = (sustain >> count) & 1;
            uint8_t nst = (uint8_t)((is_born | (is_sus & st)) & 1);
            if (col) {
                int32_t v = (int32_t)nst * 255;
                for (int c = 0; c < 3; ++c)
                    col[3 * i + c] = fade(v, col[3 * i + c], fade_in, fade_out, smooth);
            }
            fld[i] = nst;
        }
        if (hashes) hashes[s] = oracle_crc32(0, fld, n);
    }
    return 0;
}

/*
 * GameOfLifeColor (examples/game_of_life.py:112-159): a uint32 word with
 * state (1 bit), red, green, blue (8 bits each, shifts 1, 9, 17).  Every
 * buffer carries the word; absorb applies B3/S23, then each channel
 * becomes (sum over neighbours of channel+1) * new_state truncated to
 * 8 bits.  Colour: MovingAverage of state*channel.
 */
int oracle_color_life(uint32_t *fld, int32_t *col, int64_t w0, int64_t w1,
                      const int8_t *delta, const int32_t *rev,
                      int fade_in, int fade_out, int smooth, int steps, int threads) {
    lattice_t lat;
    int64_t sizes[2] = {w0, w1};
    if (setup(&lat, 2, sizes, 8, delta, rev, 0, 0)) return 1;
    set_threads(threads);
    const int64_t n = lat.n;
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t w = fld[i];
            uint32_t packed = (w & 1) + (((w >> 1) & 255) << 1) +
                              (((w >> 9) & 255) << 9) + (((w >> 17) & 255) << 17);
            for (int q = 0; q < 8; ++q) fld[i + n * (rev[q] + 1)] = packed;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t count = 0, rs = 0, gs = 0, bs = 0;
            for (int j = 0; j < 8; ++j) {
                uint32_t w = fld[neighbor_index(&lat, i, j) + n * (j + 1)];
                count += w & 1;
                rs += ((w >> 1) & 255) + 1;
                gs += ((w >> 9) & 255) + 1;
                bs += ((w >> 17) & 255) + 1;
            }
            uint32_t st = fld[i] & 1;
            uint32_t nst = (((8u >> count) & 1) | (((12u >> count) & 1) & st)) & 1;
            uint8_t r = (uint8_t)(rs * nst), g = (uint8_t)(gs * nst), b = (uint8_t)(bs * nst);
            if (col) {
                col[3 * i + 0] = fade((int32_t)(nst * r), col[3 * i + 0], fade_in, fade_out, smooth);
                col[3 * i + 1] = fade((int32_t)(nst * g), col[3 * i + 1], fade_in, fade_out, smooth);
                col[3 * i + 2] = fade((int32_t)(nst * b), col[3 * i + 2], fade_in, fade_out, smooth);
            }
            fld[i] = nst + ((uint32_t)r << 1) + ((uint32_t)g << 9) + ((uint32_t)b << 17);
        }
    }
    return 0;
}

/*
 * ShiftingSands (examples/shifting_sands.py:15-47): buffers[0] carries
 * the state, all other buffers 0; absorb sums every neighbour slot and
 * stores the sum truncated to the 1-bit property.
 */
int oracle_shifting_sands(uint8_t *fld, int32_t *col, int64_t w0, int64_t w1,
                          const int8_t *delta, const int32_t *rev,
                          int fade_in, int fade_out, int smooth, int steps, int threads) {
    lattice_t lat;
    int64_t sizes[2] = {w0, w1};
    if (setup(&lat, 2, sizes, 8, delta, rev, 0, 0)) return 1;
    set_threads(threads);
    const int64_t n = lat.n;
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint8_t st = fld[i] & 1;
            for (int q = 0; q < 8; ++q) fld[i + n * (rev[q] + 1)] = q == 0 ? st : 0;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t sum = 0;
            for (int j = 0; j < 8; ++j)
                sum += fld[neighbor_index(&lat, i, j) + n * (j + 1)] & 1;
            uint8_t nst = (uint8_t)sum & 1;
            if (col)
                for (int c = 0; c < 3; ++c)
                    col[3 * i + c] = fade((int32_t)nst * 255, col[3 * i + c], fade_in, fade_out, smooth);
            fld[i] = nst;
        }
    }
    return 0;
}

/*
 * NoiseTV (examples/noisetv.py:10-28) on a uint16 word holding the
 * RandomProperty.  Every access advances r = (r * 58321 + 11113) % 65535
 * (xentica/core/properties.py:648-655): once in absorb, once more in
 * color; the colour is (unsigned)((float)r / 65535 * 255) of the second
 * value.
 */
int oracle_noisetv(uint16_t *fld, int32_t *col, int64_t n,
                   int fade_in, int fade_out, int smooth, int steps, int threads) {
    set_threads(threads);
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t r = fld[i] & 65535u;
            r = (uint16_t)((r * 58321u + 11113u) % 65535u);
            r = (uint16_t)((r * 58321u + 11113u) % 65535u);
            if (col) {
                float u = ((float)r) / 65535;
                int32_t v = (int32_t)(unsigned int)(u * 255);
                for (int c = 0; c < 3; ++c)
                    col[3 * i + c] = fade(v, col[3 * i + c], fade_in, fade_out, smooth);
            }
            fld[i] = (uint16_t)(r & 65535u);
        }
    }
    return 0;
}

/*
 * EnergyLife (examples/energy_life.py) on one word per cell, in either of
 * the model's two layouts (the DSL packs fields in declaration order from
 * bit 0):
 *   narrow (64-bit word): energy 8 bits at 0, genome 18 at 8, mode 1 at
 *     26, xorshift32 state 32 at 27;
 *   wide (128-bit word):  energy 16 bits at 0, genome 18 at 16, mode 1 at
 *     34, xorshift32 state 32 at 35.
 * genome: birth mask bits 0-8, survival mask bits 9-17.  Emit publishes
 * energy, genome and mode (generator field 0) in every buffer slot.
 * Absorb, for each neighbour j with energy ej and mode mj: the energy
 * moves by sign(ej - e) * (|ej - e| >> (3 + m + mj)); count = number of
 * live neighbours; the richest live neighbour (first on ties, energy > 0)
 * is the parent.  The generator advances once (x ^= x << 13;
 * x ^= x >> 17; x ^= x << 5).  A live cell stays alive iff its survival
 * bit for count is set and its new energy >= sustain; a dead cell
 * becomes alive iff the parent's birth bit for count is set and its new
 * energy >= birth; a newborn takes the parent's genome, with bit
 * ((r >> 16) % 18) flipped when (r & 65535) < mutation.  Colour:
 * MovingAverage of (mode * 255, energy, (genome >> 10) & 255) of the new
 * state.  New energies stay between the neighbourhood's minimum and
 * maximum, so the narrow layout's 8 bits hold any field seeded in 0-255.
 */
typedef unsigned __int128 u128;

typedef struct {
    int wide;       /* 16-byte words (else 8-byte) */
    int e_bits;     /* energy width; genome, mode and rng follow it */
} el_layout_t;

static inline u128 el_load(const void *fld, int64_t i, const el_layout_t *lay) {
    return lay->wide ? ((const u128 *)fld)[i] : (u128)((const uint64_t *)fld)[i];
}

static inline void el_store(void *fld, int64_t i, u128 w, const el_layout_t *lay) {
    if (lay->wide)
        ((u128 *)fld)[i] = w;
    else
        ((uint64_t *)fld)[i] = (uint64_t)w;
}

static inline u128 el_pack(uint32_t e, uint32_t g, uint32_t m, uint32_t r,
                           const el_layout_t *lay) {
    const int eb = lay->e_bits;
    return (u128)(e & ((1u << eb) - 1u)) | ((u128)(g & 262143u) << eb) |
           ((u128)(m & 1u) << (eb + 18)) | ((u128)r << (eb + 19));
}

static inline void el_unpack(u128 w, const el_layout_t *lay, uint32_t *e, uint32_t *g,
                             uint32_t *m, uint32_t *r) {
    const int eb = lay->e_bits;
    *e = (uint32_t)(w & ((1u << eb) - 1u));
    *g = (uint32_t)((w >> eb) & 262143u);
    *m = (uint32_t)((w >> (eb + 18)) & 1u);
    *r = (uint32_t)(w >> (eb + 19));
}

int oracle_energy_life(void *fld, int wide, int32_t *col, int64_t w0, int64_t w1,
                       const int8_t *delta, const int32_t *rev,
                       uint32_t sustain, uint32_t birth, uint32_t mutation,
                       int fade_in, int fade_out, int smooth, int steps,
                       int threads, uint32_t *hashes) {
    lattice_t lat;
    int64_t sizes[2] = {w0, w1};
    if (setup(&lat, 2, sizes, 8, delta, rev, 0, 0)) return 1;
    set_threads(threads);
    const el_layout_t lay = {wide, wide ? 16 : 8};
    const int64_t n = lat.n;
    for (int s = 0; s < steps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t e, g, m, r;
            el_unpack(el_load(fld, i, &lay), &lay, &e, &g, &m, &r);
            const u128 b = el_pack(e, g, m, 0, &lay);
            for (int q = 0; q < 8; ++q) el_store(fld, i + n * (rev[q] + 1), b, &lay);
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t e, g, m, r;
            el_unpack(el_load(fld, i, &lay), &lay, &e, &g, &m, &r);
            uint32_t ne = e, count = 0, best_e = 0, best_g = 0;
            for (int j = 0; j < 8; ++j) {
                uint32_t ej, gj, mj, rj;
                el_unpack(el_load(fld, neighbor_index(&lat, i, j) + n * (j + 1), &lay), &lay,
                          &ej, &gj, &mj, &rj);
                uint32_t sh = 3 + m + mj;
                if (ej > e) ne += (ej - e) >> sh;
                if (e > ej) ne -= (e - ej) >> sh;
                count += mj;
                if (mj && ej > best_e) {
                    best_e = ej;
                    best_g = gj;
                }
            }
            r ^= r << 13;
            r ^= r >> 17;
            r ^= r << 5;
            uint32_t nm;
            if (m)
                nm = ((g >> (9 + count)) & 1u) && ne >= sustain;
            else
                nm = ((best_g >> count) & 1u) && best_e > 0 && ne >= birth;
            if (!m && nm) {
                g = best_g;
                if ((r & 65535u) < mutation) g ^= 1u << ((r >> 16) % 18u);
            }
            if (col) {
                const uint32_t emask = (1u << lay.e_bits) - 1u;
                col[3 * i + 0] = fade((int32_t)(nm * 255), col[3 * i + 0], fade_in, fade_out, smooth);
                col[3 * i + 1] = fade((int32_t)(ne & emask), col[3 * i + 1], fade_in, fade_out, smooth);
                col[3 * i + 2] = fade((int32_t)((g >> 10) & 255u), col[3 * i + 2], fade_in, fade_out, smooth);
            }
            el_store(fld, i, el_pack(ne, g, nm, r, &lay), &lay);
        }
        if (hashes)
            hashes[s] = oracle_crc32(0, (const uint8_t *)fld, n * (wide ? 16 : 8));
    }
    return 0;
}

/* Number of xorshift32 (13, 17, 5) steps until the state returns to
 * seed, or 0 if it does not within 2**32 steps. */
uint64_t oracle_xorshift32_period(uint32_t seed) {
    uint32_t x = seed;
    for (uint64_t t = 1; t <= 4294967296ull; ++t) {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        if (x == seed) return t;
    }
    return 0;
}

/* CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320, init and final
 * xor 0xFFFFFFFF), the checksum binascii.crc32 / zlib.crc32 compute. */
static uint32_t crc_table[256];
static int crc_ready = 0;

static void crc_init(void) {
    for (uint32_t b = 0; b < 256; ++b) {
        uint32_t c = b;
        for (int bit = 0; bit < 8; ++bit) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        crc_table[b] = c;
    }
    crc_ready = 1;
}

uint32_t oracle_crc32(uint32_t crc, const uint8_t *buf, int64_t len) {
    if (!crc_ready) crc_init();
    crc = ~crc;
    for (int64_t i = 0; i < len; ++i) crc = crc_table[(crc ^ buf[i]) & 255] ^ (crc >> 8);
    return ~crc;
}
