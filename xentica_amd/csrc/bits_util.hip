// Layout conversion (uint8 plane <-> packed bits) and chunked device CRC-32.
//
// The CRC is the reflected CRC-32 of zlib/binascii (polynomial 0xEDB88320).
// Each thread folds one chunk through a 256-entry table held in LDS with a
// zero initial register; the host combines the chunk registers with the
// GF(2) shift operator of the chunk lengths (the method of zlib's
// crc32_combine), then applies the initial value and final inversion.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xentica_hip.h"
#include "xentica_kernels.h"
#include "common.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void pack_kernel(const uint8_t *__restrict__ cells,
                                                      uint32_t *__restrict__ bits,
                                                      long long n_words) {
    const long long w = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (w >= n_words) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(cells + w * 32);
    uint32_t word = 0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint4 v = src[q];
        const uint32_t parts[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                word |= ((parts[p] >> (8 * b)) & 1u) << (q * 16 + p * 4 + b);
    }
    bits[w] = word;
}

__global__ __launch_bounds__(kBlock) void unpack_kernel(const uint32_t *__restrict__ bits,
                                                        uint8_t *__restrict__ cells,
                                                        long long n_words) {
    const long long w = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (w >= n_words) return;
    const uint32_t word = bits[w];
    uint4 *dst = reinterpret_cast<uint4 *>(cells + w * 32);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t parts[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) v |= ((word >> (q * 16 + p * 4 + b)) & 1u) << (8 * b);
            parts[p] = v;
        }
        dst[q] = make_uint4(parts[0], parts[1], parts[2], parts[3]);
    }
}

__device__ void load_crc_table(uint32_t *table) {
    for (int b = threadIdx.x; b < 256; b += blockDim.x) {
        uint32_t c = static_cast<uint32_t>(b);
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        table[b] = c;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void crc_u8_kernel(const uint8_t *__restrict__ data,
                                                        long long n, long long chunk,
                                                        long long n_chunks,
                                                        uint32_t *__restrict__ partials) {
    __shared__ uint32_t table[256];
    load_crc_table(table);
    const long long c = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_chunks) return;
    const long long lo = c * chunk;
    const long long hi = lo + chunk < n ? lo + chunk : n;
    uint32_t crc = 0;
    for (long long i = lo; i < hi; ++i) crc = table[(crc ^ data[i]) & 0xFFu] ^ (crc >> 8);
    partials[c] = crc;
}

__global__ __launch_bounds__(kBlock) void crc_bits_kernel(const uint32_t *__restrict__ bits,
                                                          long long n_cells, long long chunk,
                                                          long long n_chunks,
                                                          uint32_t *__restrict__ partials) {
    __shared__ uint32_t table[256];
    load_crc_table(table);
    const long long c = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= n_chunks) return;
    const long long lo = c * chunk;
    const long long hi = lo + chunk < n_cells ? lo + chunk : n_cells;
    uint32_t crc = 0;
    for (long long w = lo / 32; w * 32 < hi; ++w) {
        const uint32_t word = bits[w];
        const int count = hi - w * 32 < 32 ? static_cast<int>(hi - w * 32) : 32;
        for (int b = 0; b < count; ++b)
            crc = table[(crc ^ ((word >> b) & 1u)) & 0xFFu] ^ (crc >> 8);
    }
    partials[c] = crc;
}

// The sum of one bit field of every container word of a plane: field value
// (word >> shift) & (2^width - 1), width <= 32, summed in 64 bits per thread
// over a grid-stride range, then across the block (wave shuffles, then LDS)
// into partials[blockIdx.x].  Words of 16 bytes are two little-endian
// 64-bit halves.
template <typename W>
__device__ __forceinline__ uint64_t field_of(const W *words, long long i, int shift,
                                             uint32_t mask) {
    return (static_cast<uint64_t>(words[i]) >> shift) & mask;
}

template <>
__device__ __forceinline__ uint64_t field_of<ulonglong2>(const ulonglong2 *words, long long i,
                                                         int shift, uint32_t mask) {
    const ulonglong2 w = words[i];
    uint64_t v;
    if (shift >= 64)
        v = w.y >> (shift - 64);
    else if (shift == 0)
        v = w.x;
    else
        v = (w.x >> shift) | (w.y << (64 - shift));
    return v & mask;
}

template <typename W>
__global__ __launch_bounds__(kBlock) void field_sum_kernel(const W *__restrict__ words,
                                                           long long n, int shift,
                                                           uint32_t mask,
                                                           uint64_t *__restrict__ partials) {
    __shared__ uint64_t wave_sums[kBlock / 64];
    uint64_t sum = 0;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += stride)
        sum += field_of(words, i, shift, mask);
    for (int offset = 32; offset > 0; offset >>= 1) sum += __shfl_down(sum, offset, 64);
    if ((threadIdx.x & 63) == 0) wave_sums[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t total = 0;
        for (int w = 0; w < kBlock / 64; ++w) total += wave_sums[w];
        partials[blockIdx.x] = total;
    }
}

// The low byte of every int32 value (what C's and NumPy's int32 -> uint8
// casts give), four values per thread: one 16-byte load, one 4-byte store.
__global__ __launch_bounds__(kBlock) void low_bytes_kernel(const int32_t *__restrict__ src,
                                                           uint8_t *__restrict__ dst,
                                                           long long n) {
    const long long i = 4 * (blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x);
    if (i + 3 < n) {
        const int4 v = *reinterpret_cast<const int4 *>(src + i);
        *reinterpret_cast<uint32_t *>(dst + i) =
            (static_cast<uint32_t>(v.x) & 0xFFu) | ((static_cast<uint32_t>(v.y) & 0xFFu) << 8) |
            ((static_cast<uint32_t>(v.z) & 0xFFu) << 16) | (static_cast<uint32_t>(v.w) << 24);
    } else {
        for (long long k = i; k < n; ++k) dst[k] = static_cast<uint8_t>(src[k]);
    }
}

// Four cells per thread from a packed plane: their 0/1 bytes into `cells`
// (one 4-byte store) and their colours into `col` (one of two int3 values
// by state, three 16-byte stores); either output may be null.
__global__ __launch_bounds__(kBlock) void expand_kernel(const uint32_t *__restrict__ bits,
                                                        uint8_t *__restrict__ cells,
                                                        int *__restrict__ col,
                                                        long long n_quads, int4 c0a, int4 c1a) {
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= n_quads) return;
    const uint32_t nib = (bits[q >> 3] >> ((q & 7) * 4)) & 0xFu;
    if (cells)
        *reinterpret_cast<uint32_t *>(cells + 4 * q) =
            (nib & 1u) | ((nib & 2u) << 7) | ((nib & 4u) << 14) | ((nib & 8u) << 21);
    if (col) {
        int v[12];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool on = (nib >> k) & 1u;
            v[3 * k] = on ? c1a.x : c0a.x;
            v[3 * k + 1] = on ? c1a.y : c0a.y;
            v[3 * k + 2] = on ? c1a.z : c0a.z;
        }
        int4 *dst = reinterpret_cast<int4 *>(col + 12 * q);
        dst[0] = make_int4(v[0], v[1], v[2], v[3]);
        dst[1] = make_int4(v[4], v[5], v[6], v[7]);
        dst[2] = make_int4(v[8], v[9], v[10], v[11]);
    }
}

unsigned groups_for(long long items) {
    return static_cast<unsigned>((items + kBlock - 1) / kBlock);
}

}  // namespace

extern "C" {

int xen_bits_pack(const uint8_t *cells, uint32_t *bits, long long n_cells, xen_stream stream) {
    if (n_cells % 32) return xen::fail("xen_bits_pack: n_cells must be a multiple of 32", 22);
    const long long words = n_cells / 32;
    if (!words) return 0;
    hipLaunchKernelGGL(pack_kernel, dim3(groups_for(words)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), cells, bits, words);
    return xen::check(hipGetLastError(), "pack_kernel launch");
}

int xen_bits_unpack(const uint32_t *bits, uint8_t *cells, long long n_cells, xen_stream stream) {
    if (n_cells % 32) return xen::fail("xen_bits_unpack: n_cells must be a multiple of 32", 22);
    const long long words = n_cells / 32;
    if (!words) return 0;
    hipLaunchKernelGGL(unpack_kernel, dim3(groups_for(words)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), bits, cells, words);
    return xen::check(hipGetLastError(), "unpack_kernel launch");
}

int xen_bits_expand(const uint32_t *bits, uint8_t *cells, int32_t *colors, long long n_cells,
                    const int32_t *color_dead, const int32_t *color_alive, xen_stream stream) {
    if (n_cells % 32) return xen::fail("xen_bits_expand: n_cells must be a multiple of 32", 22);
    if (colors && (!color_dead || !color_alive))
        return xen::fail("xen_bits_expand: colours need both colour values", 22);
    if ((reinterpret_cast<uintptr_t>(cells) & 3) || (reinterpret_cast<uintptr_t>(colors) & 15))
        return xen::fail("xen_bits_expand: cells must be 4-byte and colors 16-byte aligned", 22);
    const long long quads = n_cells / 4;
    if (!quads || (!cells && !colors)) return 0;
    const int4 c0 = colors ? make_int4(color_dead[0], color_dead[1], color_dead[2], 0)
                           : make_int4(0, 0, 0, 0);
    const int4 c1 = colors ? make_int4(color_alive[0], color_alive[1], color_alive[2], 0)
                           : make_int4(0, 0, 0, 0);
    hipLaunchKernelGGL(expand_kernel, dim3(groups_for(quads)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), bits, cells, colors, quads, c0, c1);
    return xen::check(hipGetLastError(), "expand_kernel launch");
}

int xen_low_bytes(const int32_t *src, uint8_t *dst, long long n, xen_stream stream) {
    if (n < 0) return xen::fail("xen_low_bytes: negative count", 22);
    if (!n) return 0;
    if ((reinterpret_cast<uintptr_t>(src) & 15) || (reinterpret_cast<uintptr_t>(dst) & 3))
        return xen::fail("xen_low_bytes: src must be 16-byte and dst 4-byte aligned", 22);
    hipLaunchKernelGGL(low_bytes_kernel, dim3(groups_for((n + 3) / 4)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), src, dst, n);
    return xen::check(hipGetLastError(), "low_bytes_kernel launch");
}

int xen_crc32_u8_partials(const uint8_t *data, long long n, long long chunk, uint32_t *partials,
                          xen_stream stream) {
    if (chunk < 1) return xen::fail("xen_crc32_u8_partials: chunk must be positive", 22);
    const long long n_chunks = (n + chunk - 1) / chunk;
    if (!n_chunks) return 0;
    hipLaunchKernelGGL(crc_u8_kernel, dim3(groups_for(n_chunks)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), data, n, chunk, n_chunks,
                       partials);
    return xen::check(hipGetLastError(), "crc_u8_kernel launch");
}

int xen_crc32_bits_partials(const uint32_t *bits, long long n_cells, long long chunk,
                            uint32_t *partials, xen_stream stream) {
    if (chunk < 32 || chunk % 32)
        return xen::fail("xen_crc32_bits_partials: chunk must be a positive multiple of 32", 22);
    const long long n_chunks = (n_cells + chunk - 1) / chunk;
    if (!n_chunks) return 0;
    hipLaunchKernelGGL(crc_bits_kernel, dim3(groups_for(n_chunks)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), bits, n_cells, chunk, n_chunks,
                       partials);
    return xen::check(hipGetLastError(), "crc_bits_kernel launch");
}

int xen_field_sum_partials(const void *words, long long n, int word_bytes, int shift, int width,
                           uint64_t *partials, int n_partials, xen_stream stream) {
    if (n < 0 || n_partials < 1)
        return xen::fail("xen_field_sum_partials: negative count or no partials", 22);
    if (width < 1 || width > 32 || shift < 0 || shift + width > 8 * word_bytes)
        return xen::fail("xen_field_sum_partials: field outside the word (width 1..32)", 22);
    if (reinterpret_cast<uintptr_t>(words) % (word_bytes < 16 ? word_bytes : 16))
        return xen::fail("xen_field_sum_partials: words not aligned to their size", 22);
    const uint32_t mask = width == 32 ? 0xFFFFFFFFu : (1u << width) - 1u;
    const dim3 grid(n_partials), block(kBlock);
    const auto st = reinterpret_cast<hipStream_t>(stream);
    switch (word_bytes) {
    case 1:
        hipLaunchKernelGGL(field_sum_kernel<uint8_t>, grid, block, 0, st,
                           static_cast<const uint8_t *>(words), n, shift, mask, partials);
        break;
    case 2:
        hipLaunchKernelGGL(field_sum_kernel<uint16_t>, grid, block, 0, st,
                           static_cast<const uint16_t *>(words), n, shift, mask, partials);
        break;
    case 4:
        hipLaunchKernelGGL(field_sum_kernel<uint32_t>, grid, block, 0, st,
                           static_cast<const uint32_t *>(words), n, shift, mask, partials);
        break;
    case 8:
        hipLaunchKernelGGL(field_sum_kernel<uint64_t>, grid, block, 0, st,
                           static_cast<const uint64_t *>(words), n, shift, mask, partials);
        break;
    case 16:
        hipLaunchKernelGGL(field_sum_kernel<ulonglong2>, grid, block, 0, st,
                           static_cast<const ulonglong2 *>(words), n, shift, mask, partials);
        break;
    default:
        return xen::fail("xen_field_sum_partials: word_bytes must be 1, 2, 4, 8 or 16", 22);
    }
    return xen::check(hipGetLastError(), "field_sum_kernel launch");
}

}  // extern "C"
