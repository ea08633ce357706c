// Host check of pargo_amd/csrc/swar8.h: every byte value in every byte
// position, against per-byte scalar statements, with random bytes around.
#include <cstdio>
#include <cstdint>
#include <random>

#define __host__
#define __device__
#define __forceinline__ inline
#include "../../pargo_amd/csrc/swar8.h"

using namespace pargo;

static int fails = 0;
static void expect(bool ok, const char *what, uint32_t x, uint32_t got, uint32_t want)
{
    if (!ok && fails++ < 10) std::printf("%s(0x%08x) = 0x%08x, want 0x%08x\n", what, x, got, want);
}

int main()
{
    std::mt19937 rng(12345);
    for (int ctx = 0; ctx < 400; ctx++) {
        const uint32_t base = rng();
        for (int pos = 0; pos < 4; pos++)
            for (uint32_t v = 0; v < 256; v++) {
                const uint32_t x = (base & ~(0xFFu << (8 * pos))) | (v << (8 * pos));
                uint32_t wz = 0, wp = 0;
                for (int b = 0; b < 4; b++) {
                    const uint32_t y = (x >> (8 * b)) & 0xFF;
                    if (y == 0) wz |= 0x80u << (8 * b);
                    wp |= (uint32_t)__builtin_popcount(y) << (8 * b);
                }
                expect(swar_zero(x) == wz, "swar_zero", x, swar_zero(x), wz);
                expect(swar_popc(x) == wp, "swar_popc", x, swar_popc(x), wp);
                for (uint32_t n = 1; n <= 128; n++) {
                    uint32_t wl = 0;
                    for (int b = 0; b < 4; b++)
                        if (((x >> (8 * b)) & 0xFF) < n) wl |= 0x80u << (8 * b);
                    expect(swar_lt(x, n) == wl, "swar_lt", x, swar_lt(x, n), wl);
                }
                uint32_t fl = x & kSwarHi, wm = 0;
                for (int b = 0; b < 4; b++)
                    if ((fl >> (8 * b)) & 0x80) wm |= 0xFFu << (8 * b);
                expect(swar_mask(fl) == wm, "swar_mask", fl, swar_mask(fl), wm);
                const uint32_t hi = rng();
                for (unsigned s = 0; s < 4; s++) {
                    uint32_t ww = 0;
                    for (unsigned b = 0; b < 4; b++) {
                        const unsigned k = b + s;
                        const uint32_t byte = k < 4 ? (x >> (8 * k)) & 0xFF : (hi >> (8 * (k - 4))) & 0xFF;
                        ww |= byte << (8 * b);
                    }
                    expect(swar_window(x, hi, s) == ww, "swar_window", x, swar_window(x, hi, s), ww);
                }
            }
    }
    for (uint32_t bits = 0; bits < 16; bits++) {
        uint32_t w = 0;
        for (int b = 0; b < 4; b++)
            if ((bits >> b) & 1) w |= 0xFFu << (8 * b);
        expect(swar_bits_to_mask(bits) == w, "swar_bits_to_mask", bits, swar_bits_to_mask(bits), w);
    }
    std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails != 0;
}
