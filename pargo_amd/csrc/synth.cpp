// synth.cpp -- deterministic synthetic DEMs and IDW samples.
//
// The reference ships only a 104x124 DEM (data/dem.tif) and a 1000-point
// shapefile, so benchmark-sized inputs are generated.  Everything derives
// from SplitMix64 streams keyed by the caller's seed, so a seed fixes the
// data bit for bit on any host.
//
// DEM: value noise on an integer lattice (hashed lattice values in [0, 1),
// quintic-smoothstep bilinear interpolation), 8 octaves, base wavelength
// max(rows, cols)/4 cells, frequency x2 and amplitude x0.5 per octave,
// min-max normalised to [0, 2000] m in double and rounded to float32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/pargo.h"

namespace {

inline uint64_t splitmix64(uint64_t &state)
{
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t mix(uint64_t x)
{
    uint64_t s = x;
    return splitmix64(s);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() { return splitmix64(s); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    int uniform_int(int lo, int hi) { return lo + (int)(uniform() * (double)(hi - lo + 1)); }  // [lo, hi]
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

constexpr int kOctaves = 8;

// Hashed lattice value in [0, 1).
inline double lattice(uint64_t seed, int oct, int64_t ix, int64_t iy)
{
    uint64_t h = seed ^ 0x5041524730ull;
    h = mix(h ^ ((uint64_t)(oct + 1) * 0xD6E8FEB86659FD93ull));
    h = mix(h ^ ((uint64_t)ix * 0x9E3779B97F4A7C15ull));
    h = mix(h ^ ((uint64_t)iy * 0xC2B2AE3D27D4EB4Full));
    return (double)(h >> 11) * 0x1.0p-53;
}

inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

struct NoiseField {
    uint64_t seed;
    double base_wavelength;  // cells
    double freq(int o) const { return std::ldexp(1.0, o) / base_wavelength; }
    static double amp(int o) { return std::ldexp(1.0, -o); }
    static double amp_sum()
    {
        double s = 0;
        for (int o = 0; o < kOctaves; o++) s += amp(o);
        return s;
    }
    // Raw noise at continuous cell coordinates (x = column, y = row).
    double at(double x, double y) const
    {
        double v = 0;
        for (int o = 0; o < kOctaves; o++) {
            const double f = freq(o);
            const double xs = x * f, ys = y * f;
            const double fx0 = std::floor(xs), fy0 = std::floor(ys);
            const int64_t ix = (int64_t)fx0, iy = (int64_t)fy0;
            const double ux = fade(xs - fx0), uy = fade(ys - fy0);
            const double a = lattice(seed, o, ix, iy), b = lattice(seed, o, ix + 1, iy);
            const double c = lattice(seed, o, ix, iy + 1), d = lattice(seed, o, ix + 1, iy + 1);
            const double top = a + (b - a) * ux, bot = c + (d - c) * ux;
            v += amp(o) * (top + (bot - top) * uy);
        }
        return v;
    }
};

// One row of raw noise, reusing the lattice values along the row.
void noise_row(const NoiseField &nf, int r, int cols, double *out, std::vector<double> &latA,
               std::vector<double> &latB)
{
    std::fill(out, out + cols, 0.0);
    for (int o = 0; o < kOctaves; o++) {
        const double f = nf.freq(o);
        const double ys = r * f;
        const double fy0 = std::floor(ys);
        const int64_t iy = (int64_t)fy0;
        const double uy = fade(ys - fy0);
        const int64_t ixmax = (int64_t)std::floor((cols - 1) * f) + 1;
        latA.resize((size_t)ixmax + 1);
        latB.resize((size_t)ixmax + 1);
        for (int64_t ix = 0; ix <= ixmax; ix++) {
            latA[(size_t)ix] = lattice(nf.seed, o, ix, iy);
            latB[(size_t)ix] = lattice(nf.seed, o, ix, iy + 1);
        }
        const double a_o = NoiseField::amp(o);
        for (int c = 0; c < cols; c++) {
            const double xs = c * f;
            const double fx0 = std::floor(xs);
            const int64_t ix = (int64_t)fx0;
            const double ux = fade(xs - fx0);
            const double a = latA[(size_t)ix], b = latA[(size_t)ix + 1];
            const double cc = latB[(size_t)ix], d = latB[(size_t)ix + 1];
            const double top = a + (b - a) * ux, bot = cc + (d - cc) * ux;
            out[c] += a_o * (top + (bot - top) * uy);
        }
    }
}

}  // namespace

extern "C" {

int pargo_synth_dem(uint64_t seed, int rows, int cols, float *host_out, int nthreads)
{
    if (rows < 1 || cols < 1 || !host_out) return PARGO_E_INVALID;
    NoiseField nf{seed, std::max(rows, cols) / 4.0};
    if (nf.base_wavelength < 1.0) nf.base_wavelength = 1.0;
    const size_t n = (size_t)rows * cols;
    double gmin = INFINITY, gmax = -INFINITY;
    // Pass 1: raw noise into a double scratch row by row, keeping min/max;
    // the float output holds the raw value temporarily is not precise
    // enough, so store double rows in chunks.
    std::vector<double> raw;
    try {
        raw.resize(n);
    } catch (...) {
        return PARGO_E_NOMEM;
    }
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> la, lb;
        double lmin = INFINITY, lmax = -INFINITY;
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            double *row = raw.data() + (size_t)r * cols;
            noise_row(nf, r, cols, row, la, lb);
            for (int c = 0; c < cols; c++) {
                lmin = std::min(lmin, row[c]);
                lmax = std::max(lmax, row[c]);
            }
        }
#pragma omp critical
        {
            gmin = std::min(gmin, lmin);
            gmax = std::max(gmax, lmax);
        }
    }
    const double span = gmax - gmin;
    const double scale = span > 0 ? 2000.0 / span : 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            const size_t i = (size_t)r * cols + c;
            host_out[i] = (float)((raw[i] - gmin) * scale);
        }
    return PARGO_OK;
}

int pargo_synth_plateaus(uint64_t seed, int rows, int cols, float *dem, double fraction)
{
    if (rows < 1 || cols < 1 || !dem || !(fraction >= 0) || fraction > 0.9) return PARGO_E_INVALID;
    const size_t n = (size_t)rows * cols;
    const int64_t target = (int64_t)std::ceil(fraction * (double)n);
    std::vector<uint8_t> covered(n, 0);
    Rng rng(seed ^ 0x706C617465617531ull);
    int64_t count = 0;
    while (count < target) {
        const int w = rng.uniform_int(3, 50), h = rng.uniform_int(3, 50);
        if (w > cols || h > rows) break;
        const int r0 = rng.uniform_int(0, rows - h), c0 = rng.uniform_int(0, cols - w);
        const float v = dem[(size_t)r0 * cols + c0];
        for (int r = r0; r < r0 + h; r++)
            for (int c = c0; c < c0 + w; c++) {
                const size_t i = (size_t)r * cols + c;
                dem[i] = v;
                if (!covered[i]) {
                    covered[i] = 1;
                    count++;
                }
            }
    }
    return PARGO_OK;
}

int pargo_synth_nodata_discs(uint64_t seed, int rows, int cols, float *dem, int n_discs, double fraction,
                             float nodata, int64_t *covered_out)
{
    if (rows < 1 || cols < 1 || !dem || (n_discs <= 0 && !(fraction >= 0 && fraction <= 0.9)))
        return PARGO_E_INVALID;
    const size_t n = (size_t)rows * cols;
    const int64_t target = (int64_t)std::ceil(fraction * (double)n);
    Rng rng(seed ^ 0x6E6F646174613131ull);
    int64_t count = 0;
    for (size_t i = 0; i < n; i++)
        if (dem[i] == nodata) count++;
    int placed = 0;
    for (;;) {
        if (n_discs > 0 ? placed >= n_discs : count >= target) break;
        const int rad = rng.uniform_int(5, 40);
        const int rc = rng.uniform_int(0, rows - 1), cc = rng.uniform_int(0, cols - 1);
        for (int r = std::max(0, rc - rad); r <= std::min(rows - 1, rc + rad); r++)
            for (int c = std::max(0, cc - rad); c <= std::min(cols - 1, cc + rad); c++) {
                const int dr = r - rc, dc = c - cc;
                if (dr * dr + dc * dc > rad * rad) continue;
                const size_t i = (size_t)r * cols + c;
                if (dem[i] != nodata) {
                    dem[i] = nodata;
                    count++;
                }
            }
        placed++;
        if (placed > 100000000) break;
    }
    if (covered_out) *covered_out = count;
    return PARGO_OK;
}

int pargo_synth_samples(uint64_t seed, int n, double x0, double y0, double width, double height, double cellsize,
                        double *xs, double *ys, double *vals)
{
    if (n < 2 || !xs || !ys || !vals || !(width > 0) || !(height > 0) || !(cellsize > 0)) return PARGO_E_INVALID;
    Rng rng(seed ^ 0x73616D706C657331ull);
    // The noise field of a DEM covering the extent at this cell size.
    const double rows = height / cellsize, cols = width / cellsize;
    NoiseField nf{seed, std::max(rows, cols) / 4.0};
    const double vscale = 2000.0 / NoiseField::amp_sum();
    auto value_at = [&](double x, double y) {
        // column/row coordinates of the point, rows counted from the top
        const double c = (x - x0) / cellsize, r = (y0 + height - y) / cellsize;
        return (double)(float)(nf.at(c, r) * vscale + rng.normal());
    };
    xs[0] = (double)(float)(x0 + cellsize / 2);
    ys[0] = (double)(float)(y0 + cellsize / 2);
    xs[1] = (double)(float)(x0 + width - cellsize / 2);
    ys[1] = (double)(float)(y0 + height - cellsize / 2);
    vals[0] = value_at(xs[0], ys[0]);
    vals[1] = value_at(xs[1], ys[1]);
    auto inside = [&](double x, double y) {
        return x > xs[0] && x < xs[1] && y > ys[0] && y < ys[1];
    };
    const int n_uniform = n / 2;
    int i = 2;
    while (i < n_uniform) {
        const double x = (double)(float)(x0 + rng.uniform() * width);
        const double y = (double)(float)(y0 + rng.uniform() * height);
        if (!inside(x, y)) continue;
        xs[i] = x;
        ys[i] = y;
        vals[i] = value_at(x, y);
        i++;
    }
    double mx[10], my[10], sg[10];
    for (int k = 0; k < 10; k++) {
        mx[k] = x0 + rng.uniform() * width;
        my[k] = y0 + rng.uniform() * height;
        sg[k] = 1000.0 + rng.uniform() * 4000.0;
    }
    while (i < n) {
        const int k = std::min(9, (int)(rng.uniform() * 10.0));
        const double x = (double)(float)(mx[k] + sg[k] * rng.normal());
        const double y = (double)(float)(my[k] + sg[k] * rng.normal());
        if (!inside(x, y)) continue;  // rejection into the extent
        xs[i] = x;
        ys[i] = y;
        vals[i] = value_at(x, y);
        i++;
    }
    return PARGO_OK;
}

}  // extern "C"
