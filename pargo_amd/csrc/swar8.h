// swar8.h -- byte-parallel tests on 32-bit words: four 8-bit cells per
// lane.  Flags come back as 0x80 in the byte of each cell that passes.  The
// functions are plain integer code (host and device), checked on the host
// over every byte value by tests/test_swar8.py.
#pragma once

#include <stdint.h>

#ifndef PARGO_SWAR_FN
#define PARGO_SWAR_FN __host__ __device__ __forceinline__
#endif

namespace pargo {

constexpr uint32_t kSwarHi = 0x80808080u;
constexpr uint32_t kSwarLo7 = 0x7F7F7F7Fu;

// bytes of v repeated in all four bytes
PARGO_SWAR_FN constexpr uint32_t swar_splat(uint32_t v) { return v * 0x01010101u; }

// 0x80 in each byte of x that is zero.  (x & 0x7F) + 0x7F sets bit 7 for a
// nonzero low part without carrying into the next byte; or-ing x adds the
// bytes whose bit 7 is set.
PARGO_SWAR_FN uint32_t swar_zero(uint32_t x) { return ~(((x & kSwarLo7) + kSwarLo7) | x) & kSwarHi; }

// 0x80 in each byte of x below n (1 <= n <= 128), any byte values: the low
// seven bits plus 128 - n reach bit 7 exactly when they are >= n.
PARGO_SWAR_FN uint32_t swar_lt(uint32_t x, uint32_t n)
{
    return ~(((x & kSwarLo7) + swar_splat(128u - n)) | x) & kSwarHi;
}

// 0x80 flags to whole-byte masks (0xFF where set)
PARGO_SWAR_FN uint32_t swar_mask(uint32_t flags) { return (flags >> 7) * 0xFFu; }

// number of set bits in each byte
PARGO_SWAR_FN uint32_t swar_popc(uint32_t x)
{
    x = x - ((x >> 1) & 0x55555555u);
    x = (x & 0x33333333u) + ((x >> 2) & 0x33333333u);
    return (x + (x >> 4)) & 0x0F0F0F0Fu;
}

// four bits (cell 0 in bit 0) to whole-byte masks
PARGO_SWAR_FN uint32_t swar_bits_to_mask(uint32_t bits4)
{
    return (((bits4 & 15u) * 0x00204081u) & 0x01010101u) * 0xFFu;
}

// a where mask bytes are 0xFF, b elsewhere
PARGO_SWAR_FN uint32_t swar_select(uint32_t mask, uint32_t a, uint32_t b) { return (a & mask) | (b & ~mask); }

// four bytes starting at byte s (0..3) of the 8-byte little-endian pair
// (lo, hi): what __builtin_amdgcn_alignbyte(hi, lo, s) returns
PARGO_SWAR_FN uint32_t swar_window(uint32_t lo, uint32_t hi, unsigned s)
{
    return s == 0 ? lo : (uint32_t)((((uint64_t)hi << 32) | lo) >> (8u * s));
}

}  // namespace pargo
