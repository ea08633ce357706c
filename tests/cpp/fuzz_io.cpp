// fuzz_io.cpp -- seeded mutation fuzzing of the GeoTIFF and point-shapefile
// readers (pargo_amd/gpro/geotiff.cpp, shapefile.cpp), which read untrusted
// files.  Built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py: any out-of-bounds access, overflow or leak aborts
// the run.  A reader may accept or refuse a mutated file; it must not crash,
// must not throw, and must return dimensions it actually filled.
//
// Usage: fuzz_io <scratch dir> <iterations> <seed> <file.tif | file.shp>...
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "../../pargo_amd/gpro/geotiff.h"
#include "../../pargo_amd/gpro/shapefile.h"

namespace {

std::vector<unsigned char> slurp(const std::string &p)
{
    std::ifstream is(p, std::ios::binary);
    return std::vector<unsigned char>(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
}

void spit(const std::string &p, const std::vector<unsigned char> &b)
{
    std::ofstream os(p, std::ios::binary | std::ios::trunc);
    os.write((const char *)b.data(), (std::streamsize)b.size());
}

bool ends_with(const std::string &s, const std::string &e)
{
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

// One random edit of a file: flip bits, write an interesting value (at the
// start, where headers and tables live, half of the time), truncate, or
// duplicate a chunk.
void mutate(std::vector<unsigned char> &b, std::mt19937_64 &rng)
{
    if (b.empty()) return;
    auto pick = [&](size_t n) { return (size_t)(rng() % n); };
    const size_t head = std::min<size_t>(b.size(), 512);
    const int op = (int)(rng() % 6);
    const size_t at = (rng() & 1) ? pick(head) : pick(b.size());
    switch (op) {
    case 0:
        b[at] ^= (unsigned char)(1u << (rng() % 8));
        break;
    case 1: {
        static const uint32_t interesting[] = {0, 1, 0x7f, 0x80, 0xff, 0x7fff, 0x8000, 0xffff, 0x7fffffff,
                                               0x80000000u, 0xffffffffu, 16, 64, 4096};
        const uint32_t v = interesting[rng() % (sizeof(interesting) / sizeof(interesting[0]))];
        for (int k = 0; k < 4 && at + k < b.size(); k++) b[at + k] = (unsigned char)(v >> (8 * k));
        break;
    }
    case 2:
        b.resize(pick(b.size()) + 1);
        break;
    case 3: {
        const size_t from = pick(b.size()), len = std::min<size_t>(b.size() - from, 1 + pick(64));
        std::vector<unsigned char> chunk(b.begin() + from, b.begin() + from + len);
        b.insert(b.begin() + at, chunk.begin(), chunk.end());
        break;
    }
    case 4:
        b[at] = (unsigned char)rng();
        break;
    default: {
        const size_t n = std::min<size_t>(b.size() - at, 1 + pick(16));
        for (size_t k = 0; k < n; k++) b[at + k] = (unsigned char)rng();
        break;
    }
    }
}

int run_tif(const std::string &path)
{
    GPRO::geotiff::Info info;
    std::vector<double> data;
    std::string err;
    if (!GPRO::geotiff::read(path, info, data, err)) return 0;
    if (info.rows < 1 || info.cols < 1 || data.size() != (size_t)info.rows * (size_t)info.cols) {
        std::fprintf(stderr, "reader accepted %s with %d x %d but %zu values\n", path.c_str(), info.rows, info.cols,
                     data.size());
        return 1;
    }
    double s = 0;
    for (double v : data) s += v == v ? 0.0 : 1.0;  // touch every value
    (void)s;
    return 0;
}

int run_shp(const std::string &base)
{
    GPRO::shapefile::Points p;
    std::string err;
    if (!GPRO::shapefile::read_points(base + ".shp", 0, p, err)) return 0;
    if (p.x.size() != p.y.size() || p.x.size() != p.value.size()) {
        std::fprintf(stderr, "reader accepted %s with ragged arrays\n", base.c_str());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: fuzz_io <scratch dir> <iterations> <seed> <file.tif|file.shp>...\n");
        return 2;
    }
    const std::string tmp = argv[1];
    const long iters = std::atol(argv[2]);
    std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 10));
    int bad = 0;
    long runs = 0;
    for (int a = 4; a < argc; a++) {
        const std::string in = argv[a];
        if (ends_with(in, ".tif")) {
            const std::vector<unsigned char> orig = slurp(in);
            bad += run_tif(in);
            for (long i = 0; i < iters; i++) {
                std::vector<unsigned char> b = orig;
                const int edits = 1 + (int)(rng() % 4);
                for (int e = 0; e < edits; e++) mutate(b, rng);
                const std::string p = tmp + "/fuzz.tif";
                spit(p, b);
                bad += run_tif(p);
                runs++;
            }
        } else if (ends_with(in, ".shp")) {
            const std::string base = in.substr(0, in.size() - 4);
            const std::vector<unsigned char> shp = slurp(base + ".shp"), dbf = slurp(base + ".dbf");
            bad += run_shp(base);
            for (long i = 0; i < iters; i++) {
                std::vector<unsigned char> s = shp, d = dbf;
                const int edits = 1 + (int)(rng() % 4);
                for (int e = 0; e < edits; e++) mutate((rng() & 1) ? s : d, rng);
                const std::string p = tmp + "/fuzz";
                spit(p + ".shp", s);
                spit(p + ".dbf", d);
                bad += run_shp(p);
                runs++;
            }
        } else {
            std::fprintf(stderr, "unknown input %s\n", in.c_str());
            return 2;
        }
    }
    std::printf("fuzz_io: %ld mutated files, %d contract failures\n", runs, bad);
    return bad ? 1 : 0;
}
