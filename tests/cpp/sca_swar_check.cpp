// Host check of pargo_amd/csrc/sca_swar.h (the scad8 init pass for
// interior tiles, four cells per word) against the per-cell statements of
// the general path in sca.hip, on random tiles of every class kind.
#include <cstdio>
#include <cstdint>
#include <random>

#define __host__
#define __device__
#define __forceinline__ inline
#include "../../pargo_amd/csrc/sca_swar.h"

using namespace pargo;

static constexpr unsigned kNone = 15;
static unsigned flow_of(unsigned b) { return (b <= 8u && b != 4u) ? b : kNone; }
static unsigned excl_of(unsigned b)
{
    if (b <= 8u) return b;
    if ((b & 0xF0u) == kSwarExcl && (b & 15u) <= 8u) return b & 15u;
    return kNone;
}
static bool is_valid(unsigned b) { return b != kSwarNoData && b != kSwarInert; }

int main()
{
    std::mt19937 rng(777);
    int fails = 0;
    auto fail = [&](const char *what, int a, int b, unsigned got, unsigned want) {
        if (fails++ < 10) std::printf("%s at %d,%d: got 0x%x want 0x%x\n", what, a, b, got, want);
    };
    static SwarTile T;
    static uint8_t cls[kSwarClsRows][4 * kSwarClsWords], tgt[kSwarClsRows][4 * kSwarClsWords];
    static uint8_t inm[kSwarInmRows][4 * kSwarInmWords];
    for (int trial = 0; trial < 3000; trial++) {
        // class mix: mostly codes, with every special kind; trials differ in mix
        const int mix = trial % 4;
        for (int r = 0; r < kSwarClsRows; r++)
            for (int c = 0; c < 4 * kSwarClsWords; c++) {
                const unsigned u = rng() % 100;
                unsigned b;
                if (mix == 0 || u < 70) b = rng() % 9;
                else if (u < 76) b = kSwarExcl | (rng() % 16);
                else if (u < 82) b = kSwarNoData;
                else if (u < 86) b = kSwarInert;
                else if (u < 92) b = 255;
                else b = rng() % 256;
                cls[r][c] = (uint8_t)b;
                tgt[r][c] = (uint8_t)flow_of(b);
            }
        for (int r = 0; r < kSwarClsRows; r++)
            for (int w = 0; w < kSwarClsWords; w++) {
                uint32_t v = 0;
                for (int i = 0; i < 4; i++) v |= (uint32_t)cls[r][4 * w + i] << (8 * i);
                T.cls[r][w] = v;
                T.tgt[r][w] = swar_flow_of(v);
                for (int i = 0; i < 4; i++)
                    if (((T.tgt[r][w] >> (8 * i)) & 0xFF) != tgt[r][4 * w + i]) fail("flow_of", r, 4 * w + i, (T.tgt[r][w] >> (8 * i)) & 0xFF, tgt[r][4 * w + i]);
            }
        // inflow masks: inm column x is class column x + 4
        for (int lr = 0; lr < kSwarInmRows; lr++)
            for (int x = 0; x < 4 * kSwarInmWords; x++) {
                const int cr = lr + 1, cc = x + 4;
                const unsigned b = cls[cr][cc];
                unsigned m = 0;
                if (is_valid(b)) {
                    for (int j = 0; j < 8; j++) {
                        const int k = swar_win_of_bit(j);
                        m |= (tgt[cr + k / 3 - 1][cc + k % 3 - 1] == (unsigned)(8 - k)) ? 1u << j : 0u;
                    }
                    const unsigned ex = excl_of(b);
                    if (ex != kNone && ex != 4u) m &= ~(1u << swar_bit_of_win((int)ex));
                }
                inm[lr][x] = (uint8_t)m;
            }
        for (int lr = 0; lr < kSwarInmRows; lr++)
            for (int q = 0; q < kSwarInmWords; q++) {
                const uint32_t v = swar_inflow(T, lr, q);
                T.inm[lr][q] = v;
                for (int i = 0; i < 4; i++)
                    if (((v >> (8 * i)) & 0xFF) != inm[lr][4 * q + i]) fail("inflow", lr, 4 * q + i, (v >> (8 * i)) & 0xFF, inm[lr][4 * q + i]);
            }
        for (int conv = 0; conv < 2; conv++)
            for (int tr = 0; tr < kSwarTileR; tr++)
                for (int g = 0; g < kSwarTileC / 4; g++) {
                    const SwarCells o = swar_cells(T, tr, g, conv != 0);
                    for (int i = 0; i < 4; i++) {
                        const int tc = 4 * g + i;
                        const unsigned b = cls[tr + 2][tc + 8];
                        unsigned s = 0, npend = 0, nsrc = 0, waiting = 0;
                        if (b == kSwarInert) {
                            s = 1u << 14;
                        } else if (b != kSwarNoData) {
                            const unsigned mine = inm[tr + 1][tc + 4];
                            unsigned srcn = 0;
                            for (int j = 0; j < 8; j++) {
                                const int k = swar_win_of_bit(j);
                                srcn |= (inm[tr + 1 + k / 3 - 1][tc + 4 + k % 3 - 1] == 0 ? 1u : 0u) << j;
                            }
                            const unsigned waitable = conv ? 0xFFu : 0x0Fu;
                            nsrc = __builtin_popcount(mine & srcn);
                            npend = __builtin_popcount(mine & ~srcn & waitable);
                            const unsigned poison = (mine & ~srcn & ~waitable) != 0 ? 1u : 0u;
                            const unsigned f = flow_of(b);
                            unsigned feeds = 0;
                            if (f != kNone && mine != 0 && (conv || f > 4)) {
                                const unsigned tm = inm[tr + 1 + (int)f / 3 - 1][tc + 4 + (int)f % 3 - 1];
                                feeds = ((tm >> swar_bit_of_win(8 - (int)f)) & 1u) ? 1u : 0u;
                            }
                            s = npend | (nsrc << 4) | (poison << 8) | (feeds << 9) | ((f == kNone ? 0u : f) << 10) | (1u << 14);
                            waiting = npend ? 1u : 0u;
                        }
                        const uint32_t sw = i < 2 ? (o.s01 >> (16 * i)) & 0xFFFF : (o.s23 >> (16 * (i - 2))) & 0xFFFF;
                        if (sw != s) fail(conv ? "s (converge)" : "s", tr, tc, sw, s);
                        const bool nd = (o.nodata >> (8 * i + 7)) & 1, in = (o.inert >> (8 * i + 7)) & 1;
                        if (nd != (b == kSwarNoData)) fail("nodata flag", tr, tc, nd, b == kSwarNoData);
                        if (in != (b == kSwarInert)) fail("inert flag", tr, tc, in, b == kSwarInert);
                        if (is_valid(b)) {
                            if (((o.npend >> (8 * i)) & 0xFF) != npend) fail("npend", tr, tc, (o.npend >> (8 * i)) & 0xFF, npend);
                            if (((o.nsrc >> (8 * i)) & 0xFF) != nsrc) fail("nsrc", tr, tc, (o.nsrc >> (8 * i)) & 0xFF, nsrc);
                        }
                        if (((o.waiting >> (8 * i + 7)) & 1) != waiting) fail("waiting", tr, tc, (o.waiting >> (8 * i + 7)) & 1, waiting);
                    }
                }
    }
    std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails != 0;
}
