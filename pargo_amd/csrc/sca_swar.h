// sca_swar.h -- the scad8 init pass for interior tiles, four cells per
// 32-bit word (host and device code; the host build is checked against the
// per-cell statements of the general path by tests/test_swar8.py).
//
// A tile is 64 x 32 cells at (r0, c0).  Words hold the classes / targets of
// columns c0 - 8 .. c0 + 71 (20 words per row, rows r0 - 2 .. r0 + 33) and
// the inflow masks of columns c0 - 4 .. c0 + 67 (18 words per row, rows
// r0 - 1 .. r0 + 32).  Byte i of a word is the cell at column 4w + i of
// the row's first column.
#pragma once

#include "swar8.h"

namespace pargo {

constexpr int kSwarTileR = 32, kSwarTileC = 64;
constexpr int kSwarClsRows = kSwarTileR + 4, kSwarClsWords = 20;
constexpr int kSwarInmRows = kSwarTileR + 2, kSwarInmWords = 18;
constexpr uint32_t kSwarNoData = 254u, kSwarInert = 253u, kSwarExcl = 0x10u;

struct SwarTile {
    uint32_t cls[kSwarClsRows][kSwarClsWords];
    uint32_t tgt[kSwarClsRows][kSwarClsWords];  // target window index of each class, 15 if none
    uint32_t inm[kSwarInmRows][kSwarInmWords];  // inflow masks
};

// window position of inflow bit j (the centre, 4, has none) and back
PARGO_SWAR_FN constexpr int swar_win_of_bit(int j) { return j < 4 ? j : j + 1; }
PARGO_SWAR_FN constexpr int swar_bit_of_win(int k) { return k < 4 ? k : k - 1; }

// byte i of the result = byte sel_i of the 8-byte pool (lo: bytes 0-3, hi:
// bytes 4-7); selector 13 gives 0xFF.  The device form is one v_perm_b32.
PARGO_SWAR_FN uint32_t swar_perm(uint32_t hi, uint32_t lo, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(hi, lo, sel);
#else
    const uint64_t pool = ((uint64_t)hi << 32) | lo;
    uint32_t r = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t s = (sel >> (8 * i)) & 0xFF;
        const uint32_t byte = s < 8 ? (uint32_t)(pool >> (8 * s)) & 0xFF : (s == 12 ? 0u : 0xFFu);
        r |= byte << (8 * i);
    }
    return r;
#endif
}

// targets: the class itself for 0..8 except 4, else 15
PARGO_SWAR_FN uint32_t swar_flow_of(uint32_t b)
{
    const uint32_t ok = swar_lt(b, 9) & ~swar_zero(b ^ swar_splat(4));
    return swar_select(swar_mask(ok), b, swar_splat(15));
}

// four consecutive bytes of a row of words, starting dc bytes (-1, 0, 1)
// from word w + 1 (the word of the cells)
PARGO_SWAR_FN uint32_t swar_row_window(const uint32_t *row, int w, int dc)
{
    return dc < 0 ? swar_window(row[w], row[w + 1], 3) : (dc == 0 ? row[w + 1] : swar_window(row[w + 1], row[w + 2], 1));
}

// Inflow masks of the four cells of inflow word (lr, q) (class row lr + 1,
// class word q + 1): bit j set when the neighbour at window position
// swar_win_of_bit(j) drains into the cell and the cell does not exclude
// it; zero for NoData and inert cells.
PARGO_SWAR_FN uint32_t swar_inflow(const SwarTile &T, int lr, int q)
{
    const uint32_t b = T.cls[lr + 1][q + 1];
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k = swar_win_of_bit(j), dr = k / 3 - 1, dc = k % 3 - 1;
        const uint32_t win = swar_row_window(T.tgt[lr + 1 + dr], q, dc);
        m |= swar_zero(win ^ swar_splat((uint32_t)(8 - k))) >> (7 - j);
    }
    // the excluded window position: the class for 0..8, its low nibble for
    // PARGO_FLOW_EXCL | k (k <= 8); bit swar_bit_of_win(k) is cleared
    // (nothing for k = 4).  Masks for k = 0..7 by byte lookup, k = 8 apart.
    const uint32_t lo = b & 0x0F0F0F0Fu;
    const uint32_t isex = swar_lt(b, 9) | (swar_zero((b & 0xF0F0F0F0u) ^ swar_splat(kSwarExcl)) & swar_lt(lo, 9));
    const uint32_t exm = swar_mask(isex);
    uint32_t keep = swar_perm(0xBFDFEFFFu, 0xF7FBFDFEu, swar_select(exm, lo, swar_splat(13)));
    keep = swar_select(swar_mask(swar_zero(lo ^ swar_splat(8)) & isex), swar_splat(0x7F), keep);
    const uint32_t invalid = swar_zero(b ^ swar_splat(kSwarNoData)) | swar_zero(b ^ swar_splat(kSwarInert));
    return m & keep & ~swar_mask(invalid);
}

// The four cells of tile row tr, columns c0 + 4g ..: static words (s01:
// cells 0 and 1, s23: cells 2 and 3, 16 bits each), and per byte the
// pending and source counts and the NoData / inert flags.
struct SwarCells {
    uint32_t s01, s23;
    uint32_t npend, nsrc;     // counts per byte
    uint32_t nodata, inert;   // 0x80 flags
    uint32_t waiting;         // 0x80 flags: valid cells with npend > 0
};

PARGO_SWAR_FN SwarCells swar_cells(const SwarTile &T, int tr, int g, bool converge)
{
    const uint32_t b = T.cls[tr + 2][g + 2];
    const uint32_t f = T.tgt[tr + 2][g + 2];
    const uint32_t mine = T.inm[tr + 1][g + 1];
    uint32_t srcn = 0, feeds = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int k = swar_win_of_bit(j), dr = k / 3 - 1, dc = k % 3 - 1;
        const uint32_t win = swar_row_window(T.inm[tr + 1 + dr], g, dc);
        srcn |= swar_zero(win) >> (7 - j);
        // a cell draining to window k feeds when its target's inflow mask
        // has it (bit swar_bit_of_win(8 - k) of that mask), and then only in
        // the converging mode or when the target comes later in the sweep
        if (converge || k > 4) {
            const int tb = swar_bit_of_win(8 - k);
            feeds |= swar_zero(f ^ swar_splat((uint32_t)k)) & (win << (7 - tb));
        }
    }
    feeds &= ~swar_zero(mine);  // sources never walk
    const uint32_t waitable = converge ? 0xFFFFFFFFu : 0x0F0F0F0Fu;  // window bits 0-3: NW, N, NE, W
    SwarCells o;
    o.nsrc = swar_popc(mine & srcn);
    o.npend = swar_popc(mine & ~srcn & waitable);
    const uint32_t poison = ~swar_zero(mine & ~srcn & ~waitable) & kSwarHi;
    const uint32_t fz = f & ~swar_mask(swar_zero(f ^ swar_splat(15)));
    o.nodata = swar_zero(b ^ swar_splat(kSwarNoData));
    o.inert = swar_zero(b ^ swar_splat(kSwarInert));
    const uint32_t vm = ~swar_mask(o.nodata | o.inert);
    const uint32_t lb = (o.npend | (o.nsrc << 4)) & vm;
    const uint32_t hb = (((poison >> 7) | ((feeds >> 7) << 1) | (fz << 2) | swar_splat(0x40)) & vm) |
                        (swar_splat(0x40) & swar_mask(o.inert));
    o.s01 = swar_perm(hb, lb, 0x05010400u);
    o.s23 = swar_perm(hb, lb, 0x07030602u);
    o.waiting = ~swar_zero(o.npend) & kSwarHi & vm;
    return o;
}

}  // namespace pargo
