// geotiff.cpp -- see geotiff.h.
#include "geotiff.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>

#include "../csrc/faults.h"

namespace GPRO {
namespace geotiff {

namespace {

enum : uint16_t {
    T_WIDTH = 256, T_LENGTH = 257, T_BITS = 258, T_COMPRESSION = 259, T_PHOTOMETRIC = 262,
    T_STRIP_OFFSETS = 273, T_SPP = 277, T_ROWS_PER_STRIP = 278, T_STRIP_BYTES = 279, T_PLANAR = 284,
    T_PREDICTOR = 317, T_TILE_W = 322, T_TILE_H = 323, T_TILE_OFFSETS = 324, T_TILE_BYTES = 325,
    T_SAMPLE_FORMAT = 339, T_PIXEL_SCALE = 33550, T_TIEPOINT = 33922, T_TRANSFORM = 34264,
    T_GEOKEYS = 34735, T_GEODOUBLES = 34736, T_GEOASCII = 34737, T_GDAL_NODATA = 42113,
};

struct Entry {
    uint16_t type = 0;
    uint64_t count = 0;
    uint64_t value_off = 0;  // file offset of the values
};

class File {
  public:
    std::vector<unsigned char> b;
    bool le = true, big = false;

    bool ok(uint64_t off, uint64_t n) const { return off <= b.size() && n <= b.size() - off; }
    uint64_t u(uint64_t off, int n) const
    {
        uint64_t v = 0;
        for (int i = 0; i < n; i++) {
            const uint64_t byte = b[off + (le ? i : n - 1 - i)];
            v |= byte << (8 * i);
        }
        return v;
    }
    double f64(uint64_t off) const
    {
        uint64_t v = u(off, 8);
        double d;
        std::memcpy(&d, &v, 8);
        return d;
    }
    float f32(uint64_t off) const
    {
        uint32_t v = (uint32_t)u(off, 4);
        float f;
        std::memcpy(&f, &v, 4);
        return f;
    }
};

int type_size(uint16_t t)
{
    switch (t) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
    }
}

// count * size of an entry's values, or false when it cannot fit the file
// (the product would overflow only for counts no file can hold)
bool entry_bytes(const File &f, const Entry &e, uint64_t &bytes)
{
    const int sz = type_size(e.type);
    if (!sz || e.count > f.b.size()) return false;
    bytes = e.count * (uint64_t)sz;
    return f.ok(e.value_off, bytes);
}

bool values(const File &f, const Entry &e, std::vector<double> &out)
{
    const int sz = type_size(e.type);
    uint64_t bytes = 0;
    if (!entry_bytes(f, e, bytes)) return false;
    out.resize(e.count);
    for (uint64_t i = 0; i < e.count; i++) {
        const uint64_t o = e.value_off + i * sz;
        switch (e.type) {
        case 1: case 7: case 2: out[i] = (double)f.b[o]; break;
        case 6: out[i] = (double)(int8_t)f.b[o]; break;
        case 3: out[i] = (double)f.u(o, 2); break;
        case 8: out[i] = (double)(int16_t)f.u(o, 2); break;
        case 4: case 13: out[i] = (double)f.u(o, 4); break;
        case 9: out[i] = (double)(int32_t)f.u(o, 4); break;
        case 16: case 18: out[i] = (double)f.u(o, 8); break;
        case 17: out[i] = (double)(int64_t)f.u(o, 8); break;
        case 11: out[i] = (double)f.f32(o); break;
        case 12: out[i] = f.f64(o); break;
        case 5: out[i] = (double)f.u(o, 4) / (double)f.u(o + 4, 4); break;
        case 10: out[i] = (double)(int32_t)f.u(o, 4) / (double)(int32_t)f.u(o + 4, 4); break;
        }
    }
    return true;
}

// offsets/byte counts may exceed 2^53 in principle; keep them integral
bool uvalues(const File &f, const Entry &e, std::vector<uint64_t> &out)
{
    const int sz = type_size(e.type);
    uint64_t bytes = 0;
    if (!entry_bytes(f, e, bytes)) return false;
    out.resize(e.count);
    for (uint64_t i = 0; i < e.count; i++) out[i] = f.u(e.value_off + i * sz, sz);
    return true;
}

std::string ascii(const File &f, const Entry &e)
{
    if (!f.ok(e.value_off, e.count)) return "";
    std::string s((const char *)&f.b[e.value_off], (size_t)e.count);
    while (!s.empty() && (s.back() == '\0' || s.back() == '|')) s.pop_back();
    return s;
}

// TIFF LZW: MSB-first codes, 9..12 bits, clear 256, end 257, the code width
// grows when the next free code reaches 511, 1023 and 2047.
bool lzw_decode(const unsigned char *in, size_t n, std::vector<unsigned char> &out, size_t expect)
{
    std::vector<int> prefix(4096, -1);
    std::vector<unsigned char> last(4096), first(4096);
    std::vector<int> length(4096, 1);
    for (int i = 0; i < 256; i++) {
        last[i] = first[i] = (unsigned char)i;
        length[i] = 1;
    }
    out.clear();
    out.reserve(expect);
    size_t bitpos = 0;
    int width = 9, next = 258, old = -1;
    std::vector<unsigned char> tmp;
    auto emit = [&](int code) {
        const int len = length[code];
        const size_t base = out.size();
        out.resize(base + len);
        int c = code;
        for (int k = len - 1; k >= 0; k--) {
            out[base + k] = last[c];
            c = prefix[c];
        }
    };
    while (true) {
        if (bitpos + width > n * 8) break;
        int code = 0;
        for (int k = 0; k < width; k++) {
            const size_t bit = bitpos + k;
            code = (code << 1) | ((in[bit >> 3] >> (7 - (bit & 7))) & 1);
        }
        bitpos += width;
        if (code == 257) break;
        if (code == 256) {
            next = 258;
            width = 9;
            old = -1;
            continue;
        }
        if (old < 0) {
            if (code > 255) return false;
            emit(code);
            old = code;
            continue;
        }
        if (next >= 4096) return false;
        if (code < next) {
            emit(code);
            prefix[next] = old;
            last[next] = first[code];
            first[next] = first[old];
            length[next] = length[old] + 1;
        } else if (code == next) {
            prefix[next] = old;
            last[next] = first[old];
            first[next] = first[old];
            length[next] = length[old] + 1;
            emit(code);
        } else {
            return false;
        }
        next++;
        if (next == 511) width = 10;
        else if (next == 1023) width = 11;
        else if (next == 2047) width = 12;
        old = code;
        if (out.size() >= expect) break;
    }
    return true;
}

bool packbits_decode(const unsigned char *in, size_t n, std::vector<unsigned char> &out, size_t expect)
{
    out.clear();
    size_t i = 0;
    while (i < n && out.size() < expect) {
        const int8_t h = (int8_t)in[i++];
        if (h >= 0) {
            const size_t cnt = (size_t)h + 1;
            if (i + cnt > n) return false;
            out.insert(out.end(), in + i, in + i + cnt);
            i += cnt;
        } else if (h != -128) {
            if (i >= n) return false;
            out.insert(out.end(), (size_t)(1 - h), in[i++]);
        }
    }
    return true;
}

bool parse(const std::string &path, File &f, std::map<uint16_t, Entry> &tags, std::string &err)
{
    std::ifstream is(path, std::ios::binary);
    if (!is) {
        err = "cannot open " + path;
        return false;
    }
    pargo::alloc_point("geotiff.file");
    f.b.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
    if (f.b.size() < 8) {
        err = path + ": not a TIFF file";
        return false;
    }
    if (f.b[0] == 'I' && f.b[1] == 'I') f.le = true;
    else if (f.b[0] == 'M' && f.b[1] == 'M') f.le = false;
    else {
        err = path + ": not a TIFF file";
        return false;
    }
    const uint64_t magic = f.u(2, 2);
    uint64_t ifd;
    if (magic == 42) {
        ifd = f.u(4, 4);
    } else if (magic == 43) {
        f.big = true;
        if (!f.ok(0, 16)) return false;
        ifd = f.u(8, 8);
    } else {
        err = path + ": not a TIFF file";
        return false;
    }
    const int cnt_sz = f.big ? 8 : 2, ent_sz = f.big ? 20 : 12, val_sz = f.big ? 8 : 4;
    if (!f.ok(ifd, cnt_sz)) {
        err = path + ": bad IFD offset";
        return false;
    }
    const uint64_t n = f.u(ifd, cnt_sz);
    if (n > f.b.size() / ent_sz || !f.ok(ifd + cnt_sz, n * ent_sz)) {
        err = path + ": truncated IFD";
        return false;
    }
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t o = ifd + cnt_sz + i * ent_sz;
        Entry e;
        const uint16_t tag = (uint16_t)f.u(o, 2);
        e.type = (uint16_t)f.u(o + 2, 2);
        e.count = f.u(o + 4, f.big ? 8 : 4);
        const uint64_t vo = o + (f.big ? 12 : 8);
        // values that fit the entry's value field sit there, others at an offset
        const bool inline_vals = e.count <= (uint64_t)val_sz && e.count * (uint64_t)type_size(e.type) <= (uint64_t)val_sz;
        e.value_off = inline_vals ? vo : f.u(vo, val_sz);
        tags[tag] = e;
    }
    return true;
}

double get1(const File &f, const std::map<uint16_t, Entry> &tags, uint16_t tag, double dflt)
{
    auto it = tags.find(tag);
    if (it == tags.end()) return dflt;
    std::vector<double> v;
    if (!values(f, it->second, v) || v.empty()) return dflt;
    return v[0];
}

// An integer tag within [lo, hi]; false when it is not (a fractional, huge
// or non-finite value never reaches an int conversion).
bool get_int(const File &f, const std::map<uint16_t, Entry> &tags, uint16_t tag, long long dflt, long long lo,
             long long hi, long long &out)
{
    const double v = get1(f, tags, tag, (double)dflt);
    if (!(v >= (double)lo && v <= (double)hi) || v != std::floor(v)) return false;
    out = (long long)v;
    return true;
}

bool fill_info(const File &f, const std::map<uint16_t, Entry> &tags, Info &info, std::string &err,
               const std::string &path)
{
    long long cols = 0, rows = 0, bits = 1, fmt = 1;
    if (!get_int(f, tags, T_WIDTH, 0, 1, INT32_MAX, cols) || !get_int(f, tags, T_LENGTH, 0, 1, INT32_MAX, rows)) {
        err = path + ": missing or bad image dimensions";
        return false;
    }
    if (!get_int(f, tags, T_BITS, 1, 1, 64, bits) || !get_int(f, tags, T_SAMPLE_FORMAT, 1, 1, 6, fmt)) {
        err = path + ": bad sample description";
        return false;
    }
    info.cols = (int)cols;
    info.rows = (int)rows;
    info.bits = (int)bits;
    info.sample_format = (int)fmt;
    std::vector<double> scale, tie, m;
    auto it = tags.find(T_PIXEL_SCALE);
    if (it != tags.end()) values(f, it->second, scale);
    it = tags.find(T_TIEPOINT);
    if (it != tags.end()) values(f, it->second, tie);
    it = tags.find(T_TRANSFORM);
    if (it != tags.end()) values(f, it->second, m);
    if (scale.size() >= 2 && tie.size() >= 6) {
        info.transform[1] = scale[0];
        info.transform[5] = -scale[1];
        info.transform[2] = info.transform[4] = 0;
        info.transform[0] = tie[3] - tie[0] * scale[0];
        info.transform[3] = tie[4] + tie[1] * scale[1];
    } else if (m.size() >= 16) {
        info.transform[0] = m[3];
        info.transform[1] = m[0];
        info.transform[2] = m[1];
        info.transform[3] = m[7];
        info.transform[4] = m[4];
        info.transform[5] = m[5];
    }
    it = tags.find(T_GEOKEYS);
    if (it != tags.end()) {
        std::vector<double> v;
        values(f, it->second, v);
        info.geokeys.assign(v.begin(), v.end());
    }
    it = tags.find(T_GEODOUBLES);
    if (it != tags.end()) values(f, it->second, info.geodoubles);
    it = tags.find(T_GEOASCII);
    if (it != tags.end()) info.geoascii = ascii(f, it->second);
    it = tags.find(T_GDAL_NODATA);
    if (it != tags.end()) {
        info.nodata_text = ascii(f, it->second);
        char *end = nullptr;
        const double v = std::strtod(info.nodata_text.c_str(), &end);
        if (end != info.nodata_text.c_str()) {
            info.has_nodata = true;
            info.nodata = v;
        }
    }
    return true;
}

}  // namespace

bool read_info(const std::string &path, Info &info, std::string &err)
{
    File f;
    std::map<uint16_t, Entry> tags;
    if (!parse(path, f, tags, err)) return false;
    return fill_info(f, tags, info, err, path);
}

bool read(const std::string &path, Info &info, std::vector<double> &data, std::string &err)
{
    File f;
    std::map<uint16_t, Entry> tags;
    if (!parse(path, f, tags, err)) return false;
    if (!fill_info(f, tags, info, err, path)) return false;
    long long spp = 1, planar = 1, comp = 1, pred = 1;
    if (!get_int(f, tags, T_SPP, 1, 1, 65535, spp) || !get_int(f, tags, T_PLANAR, 1, 1, 2, planar) ||
        !get_int(f, tags, T_COMPRESSION, 1, 1, 65535, comp) || !get_int(f, tags, T_PREDICTOR, 1, 1, 3, pred)) {
        err = path + ": bad sample layout, compression or predictor tag";
        return false;
    }
    const int bps = info.bits / 8;
    const int fmt = info.sample_format;
    if (!(bps == 1 || bps == 2 || bps == 4 || bps == 8) || info.bits % 8) {
        err = path + ": unsupported bits per sample";
        return false;
    }
    if (comp != 1 && comp != 5 && comp != 32773) {
        err = path + ": unsupported compression " + std::to_string(comp);
        return false;
    }
    const bool tiled = tags.count(T_TILE_OFFSETS) != 0;
    long long bw = info.cols, bh = info.rows;
    bool block_ok = true;
    if (tiled) {
        block_ok = get_int(f, tags, T_TILE_W, 0, 1, INT32_MAX, bw) && get_int(f, tags, T_TILE_H, 0, 1, INT32_MAX, bh);
    } else {
        // RowsPerStrip may be 2^32-1 ("one strip"); only its minimum with the rows matters
        const double rps = get1(f, tags, T_ROWS_PER_STRIP, info.rows);
        block_ok = rps >= 1 && rps == std::floor(rps);
        if (block_ok) bh = (long long)std::min<double>(rps, info.rows);
    }
    if (!block_ok) {
        err = path + ": bad block size";
        return false;
    }
    std::vector<uint64_t> offs, counts;
    if (!uvalues(f, tags[tiled ? T_TILE_OFFSETS : T_STRIP_OFFSETS], offs) ||
        !uvalues(f, tags[tiled ? T_TILE_BYTES : T_STRIP_BYTES], counts) || offs.size() != counts.size()) {
        err = path + ": bad strip/tile tables";
        return false;
    }
    const int sample_stride = planar == 1 ? (int)spp : 1;  // band 1 only
    const long long across = (info.cols + bw - 1) / bw, down = (info.rows + bh - 1) / bh;
    if ((uint64_t)across * (uint64_t)down > offs.size()) {
        err = path + ": too few strips/tiles";
        return false;
    }
    // Sizes the file cannot back are refused before anything is allocated:
    // a block decodes to at most 3641 bytes per compressed byte (a 9-bit LZW
    // code emits at most 4096 bytes; PackBits at most 64 per byte), the
    // image to at most that over the whole file, and an uncompressed block
    // must be present in full.  Tiles may overhang the image (a 10 x 10
    // image in one 256 x 256 tile), so a block may be 64 MiB even then.
    const uint64_t expand = comp == 1 ? 1 : 3641;
    const uint64_t image_bytes = (uint64_t)info.rows * (uint64_t)info.cols * (uint64_t)bps * (uint64_t)sample_stride;
    const uint64_t block_px = (uint64_t)bw * (uint64_t)bh;
    if (image_bytes / expand > f.b.size() || block_px > (1ull << 40) ||
        block_px * (uint64_t)bps * (uint64_t)sample_stride > std::max<uint64_t>(image_bytes, 1ull << 26)) {
        err = path + ": the image or block size exceeds what the file holds";
        return false;
    }
    pargo::alloc_point("geotiff.raster");
    data.assign((size_t)info.rows * info.cols, 0.0);
    pargo::alloc_point("geotiff.block");
    std::vector<unsigned char> buf;
    const size_t block_bytes = (size_t)block_px * bps * sample_stride;
    for (long long by = 0; by < down; by++)
        for (long long bx = 0; bx < across; bx++) {
            const size_t k = (size_t)by * across + bx;
            if (!f.ok(offs[k], counts[k])) {
                err = path + ": truncated data";
                return false;
            }
            const unsigned char *src = f.b.data() + offs[k];
            if (comp == 1) {
                // rows of the block inside the image must be present
                const uint64_t need = (uint64_t)std::min<long long>(bh, info.rows - (long long)by * bh) * bw * bps *
                                      sample_stride;
                if (counts[k] < need) {
                    err = path + ": truncated strip or tile";
                    return false;
                }
                buf.assign(src, src + std::min<uint64_t>(counts[k], block_bytes));
            } else if (comp == 5) {
                if (!lzw_decode(src, counts[k], buf, block_bytes)) {
                    err = path + ": corrupt LZW data";
                    return false;
                }
            } else if (!packbits_decode(src, counts[k], buf, block_bytes)) {
                err = path + ": corrupt PackBits data";
                return false;
            }
            buf.resize(block_bytes, 0);
            const size_t row_bytes = (size_t)bw * bps * sample_stride;
            if (pred == 2) {  // horizontal differencing, per sample
                for (int y = 0; y < bh; y++) {
                    unsigned char *rb = buf.data() + y * row_bytes;
                    for (long long x = sample_stride; x < bw * sample_stride; x++) {
                        const size_t a = (size_t)x * bps, p = (size_t)(x - sample_stride) * bps;
                        uint64_t cur = 0, prev = 0;
                        for (int q = 0; q < bps; q++) {
                            const int sh = 8 * (f.le ? q : bps - 1 - q);
                            cur |= (uint64_t)rb[a + q] << sh;
                            prev |= (uint64_t)rb[p + q] << sh;
                        }
                        cur += prev;
                        for (int q = 0; q < bps; q++) rb[a + q] = (unsigned char)(cur >> (8 * (f.le ? q : bps - 1 - q)));
                    }
                }
            } else if (pred == 3) {  // floating-point predictor: bytes split by significance
                std::vector<unsigned char> row(row_bytes);
                for (int y = 0; y < bh; y++) {
                    unsigned char *rb = buf.data() + y * row_bytes;
                    for (size_t x = 1; x < row_bytes; x++) rb[x] = (unsigned char)(rb[x] + rb[x - 1]);
                    const size_t wn = (size_t)bw * sample_stride;
                    for (size_t x = 0; x < wn; x++)
                        for (int q = 0; q < bps; q++) {
                            // byte q of the big-endian value sits in plane q
                            const unsigned char v = rb[(size_t)q * wn + x];
                            row[x * bps + (f.le ? bps - 1 - q : q)] = v;
                        }
                    std::memcpy(rb, row.data(), row_bytes);
                }
            }
            for (int y = 0; y < bh; y++) {
                const long long r = by * bh + y;
                if (r >= info.rows) break;
                for (int x = 0; x < bw; x++) {
                    const long long c = bx * bw + x;
                    if (c >= info.cols) break;
                    const size_t o = ((size_t)y * bw + x) * bps * sample_stride;
                    uint64_t raw = 0;
                    for (int q = 0; q < bps; q++) raw |= (uint64_t)buf[o + q] << (8 * (f.le ? q : bps - 1 - q));
                    double v;
                    if (fmt == 3 && bps == 4) {
                        float fv;
                        uint32_t r32 = (uint32_t)raw;
                        std::memcpy(&fv, &r32, 4);
                        v = (double)fv;
                    } else if (fmt == 3 && bps == 8) {
                        std::memcpy(&v, &raw, 8);
                    } else if (fmt == 2) {
                        const int sh = 64 - 8 * bps;
                        v = (double)((int64_t)(raw << sh) >> sh);
                    } else {
                        v = (double)raw;
                    }
                    data[(size_t)r * info.cols + c] = v;
                }
            }
        }
    return true;
}

namespace {

struct Writer {
    std::vector<unsigned char> b;
    void u16(uint16_t v) { b.push_back(v & 0xff), b.push_back(v >> 8); }
    void u32(uint32_t v)
    {
        for (int i = 0; i < 4; i++) b.push_back((v >> (8 * i)) & 0xff);
    }
    void f64(double d)
    {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        for (int i = 0; i < 8; i++) b.push_back((v >> (8 * i)) & 0xff);
    }
    void align() { if (b.size() & 1) b.push_back(0); }
};

}  // namespace

bool write(const std::string &path, const Info &info, const double *data, Type type, std::string &err)
{
    const int bps = type == Type::Float64 ? 8 : (type == Type::UInt8 ? 1 : 4);
    const uint64_t image_bytes = (uint64_t)info.rows * info.cols * bps;
    if (image_bytes + (uint64_t)info.rows * 8 + 65536 >= (1ull << 32)) {
        err = path + ": raster too large for a classic TIFF";
        return false;
    }
    std::ofstream os(path, std::ios::binary | std::ios::trunc);
    if (!os) {
        err = "cannot create " + path;
        return false;
    }
    pargo::alloc_point("geotiff.header");
    Writer w;
    w.b.reserve(1024);
    w.b.push_back('I'), w.b.push_back('I');
    w.u16(42);
    w.u32(0);  // IFD offset, patched below
    // metadata arrays first, pixel rows streamed after them
    struct Tag {
        uint16_t tag, type;
        uint32_t count, value;
    };
    std::vector<Tag> tags;
    const uint32_t rows = (uint32_t)info.rows, cols = (uint32_t)info.cols;
    // layout: [header][strip offsets][strip counts][scale][tiepoint][geo*][nodata][IFD][pixels]
    const uint32_t off_offsets = (uint32_t)w.b.size();
    w.b.resize(w.b.size() + 4 * (size_t)rows);
    const uint32_t off_counts = (uint32_t)w.b.size();
    for (uint32_t r = 0; r < rows; r++) w.u32(cols * bps);
    const uint32_t off_scale = (uint32_t)w.b.size();
    w.f64(info.transform[1]), w.f64(-info.transform[5]), w.f64(0.0);
    const uint32_t off_tie = (uint32_t)w.b.size();
    w.f64(0.0), w.f64(0.0), w.f64(0.0), w.f64(info.transform[0]), w.f64(info.transform[3]), w.f64(0.0);
    uint32_t off_keys = 0, off_dbl = 0, off_ascii = 0, off_nd = 0;
    if (!info.geokeys.empty()) {
        off_keys = (uint32_t)w.b.size();
        for (uint16_t k : info.geokeys) w.u16(k);
        w.align();
    }
    if (!info.geodoubles.empty()) {
        off_dbl = (uint32_t)w.b.size();
        for (double d : info.geodoubles) w.f64(d);
    }
    std::string ga = info.geoascii;
    if (!ga.empty()) {
        if (ga.back() != '|') ga += '|';
        off_ascii = (uint32_t)w.b.size();
        w.b.insert(w.b.end(), ga.begin(), ga.end());
        w.b.push_back(0);
        w.align();
    }
    std::string nd;
    if (info.has_nodata) {
        char tmp[64];
        std::snprintf(tmp, sizeof(tmp), "%.17g", info.nodata);
        nd = tmp;
        off_nd = (uint32_t)w.b.size();
        w.b.insert(w.b.end(), nd.begin(), nd.end());
        w.b.push_back(0);
        w.align();
    }
    const uint16_t fmt = type == Type::UInt8 ? 1 : (type == Type::Int32 ? 2 : 3);
    tags.push_back({T_WIDTH, 4, 1, cols});
    tags.push_back({T_LENGTH, 4, 1, rows});
    tags.push_back({T_BITS, 3, 1, (uint32_t)(8 * bps)});
    tags.push_back({T_COMPRESSION, 3, 1, 1});
    tags.push_back({T_PHOTOMETRIC, 3, 1, 1});
    tags.push_back({T_STRIP_OFFSETS, 4, rows, rows == 1 ? 0u : off_offsets});
    tags.push_back({T_SPP, 3, 1, 1});
    tags.push_back({T_ROWS_PER_STRIP, 4, 1, 1});
    tags.push_back({T_STRIP_BYTES, 4, rows, rows == 1 ? cols * bps : off_counts});
    tags.push_back({T_PLANAR, 3, 1, 1});
    tags.push_back({T_SAMPLE_FORMAT, 3, 1, fmt});
    tags.push_back({T_PIXEL_SCALE, 12, 3, off_scale});
    tags.push_back({T_TIEPOINT, 12, 6, off_tie});
    if (off_keys) tags.push_back({T_GEOKEYS, 3, (uint32_t)info.geokeys.size(), off_keys});
    if (off_dbl) tags.push_back({T_GEODOUBLES, 12, (uint32_t)info.geodoubles.size(), off_dbl});
    if (off_ascii) tags.push_back({T_GEOASCII, 2, (uint32_t)ga.size() + 1, off_ascii});
    if (off_nd) {
        if (nd.size() + 1 <= 4) {
            uint32_t v = 0;
            std::memcpy(&v, nd.c_str(), nd.size());
            tags.push_back({T_GDAL_NODATA, 2, (uint32_t)nd.size() + 1, v});
        } else {
            tags.push_back({T_GDAL_NODATA, 2, (uint32_t)nd.size() + 1, off_nd});
        }
    }
    // IFD
    const uint32_t ifd = (uint32_t)w.b.size();
    w.u16((uint16_t)tags.size());
    size_t pos_first_strip = 0;
    for (const Tag &t : tags) {
        w.u16(t.tag);
        w.u16(t.type);
        w.u32(t.count);
        if (t.tag == T_STRIP_OFFSETS && rows == 1) pos_first_strip = w.b.size();
        if (t.type == 3 && t.count == 1) {
            w.u16((uint16_t)t.value);
            w.u16(0);
        } else {
            w.u32(t.value);
        }
    }
    w.u32(0);
    const uint32_t pixels = (uint32_t)w.b.size();
    // patch header and strip offsets
    for (int i = 0; i < 4; i++) w.b[4 + i] = (ifd >> (8 * i)) & 0xff;
    if (rows == 1) {
        for (int i = 0; i < 4; i++) w.b[pos_first_strip + i] = (pixels >> (8 * i)) & 0xff;
    } else {
        for (uint32_t r = 0; r < rows; r++) {
            const uint32_t o = pixels + r * cols * bps;
            for (int i = 0; i < 4; i++) w.b[off_offsets + 4 * r + i] = (o >> (8 * i)) & 0xff;
        }
    }
    os.write((const char *)w.b.data(), (std::streamsize)w.b.size());
    pargo::alloc_point("geotiff.row");
    std::vector<unsigned char> row((size_t)cols * bps);
    for (uint32_t r = 0; r < rows; r++) {
        const double *src = data + (size_t)r * cols;
        for (uint32_t c = 0; c < cols; c++) {
            if (type == Type::Float64) {
                std::memcpy(&row[(size_t)c * 8], &src[c], 8);
            } else if (type == Type::Float32) {
                const float v = (float)src[c];
                std::memcpy(&row[(size_t)c * 4], &v, 4);
            } else if (type == Type::Int32) {
                const double v = src[c];
                const int32_t i = v != v ? 0 : (v < -2147483648.0 ? INT32_MIN : (v > 2147483647.0 ? INT32_MAX : (int32_t)v));
                std::memcpy(&row[(size_t)c * 4], &i, 4);
            } else {
                const double v = src[c];
                row[c] = (unsigned char)(v != v || v < 0 ? 0 : (v > 255 ? 255 : v));
            }
        }
        os.write((const char *)row.data(), (std::streamsize)row.size());
    }
    if (!os) {
        err = "error writing " + path;
        return false;
    }
    return true;
}

}  // namespace geotiff
}  // namespace GPRO
