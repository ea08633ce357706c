// shapefile.cpp -- see shapefile.h.
#include "shapefile.h"

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

#include "../csrc/faults.h"

namespace GPRO {
namespace shapefile {

namespace {

bool slurp(const std::string &path, std::vector<unsigned char> &b)
{
    std::ifstream is(path, std::ios::binary);
    if (!is) return false;
    pargo::alloc_point("shapefile.file");
    b.assign(std::istreambuf_iterator<char>(is), std::istreambuf_iterator<char>());
    return true;
}

int32_t be32(const unsigned char *p) { return (int32_t)((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]); }
int32_t le32(const unsigned char *p) { return (int32_t)((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]); }
uint16_t le16(const unsigned char *p) { return (uint16_t)(p[0] | p[1] << 8); }
double led(const unsigned char *p)
{
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = v << 8 | p[i];
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

std::string base_of(const std::string &shp)
{
    if (shp.size() > 4) {
        std::string ext = shp.substr(shp.size() - 4);
        for (char &c : ext) c = (char)std::tolower((unsigned char)c);
        if (ext == ".shp") return shp.substr(0, shp.size() - 4);
    }
    return shp;
}

}  // namespace

bool read_points(const std::string &path, int field_index, Points &out, std::string &err)
{
    const std::string base = base_of(path);
    std::vector<unsigned char> shp, dbf, prj;
    if (!slurp(base + ".shp", shp) || shp.size() < 100 || be32(&shp[0]) != 9994) {
        err = "[ERROR] Open failed: " + path;
        return false;
    }
    if (!slurp(base + ".dbf", dbf) || dbf.size() < 32) {
        err = "[ERROR] Open failed: " + base + ".dbf";
        return false;
    }
    if (slurp(base + ".prj", prj)) out.projection_wkt.assign(prj.begin(), prj.end());
    // DBF header and field descriptors
    const uint32_t nrec = (uint32_t)le32(&dbf[4]);
    const uint16_t hlen = le16(&dbf[8]), rlen = le16(&dbf[10]);
    if (hlen < 33 || rlen < 1) {
        err = "bad DBF header in " + base + ".dbf";
        return false;
    }
    struct Field {
        int off, len;
        char type;
    };
    std::vector<Field> fields;
    int off = 1;  // deletion flag
    for (size_t o = 32; o + 32 <= dbf.size() && dbf[o] != 0x0D; o += 32) {
        char name[12] = {0};
        std::memcpy(name, &dbf[o], 11);
        out.field_names.push_back(name);
        fields.push_back({off, dbf[o + 16], (char)dbf[o + 11]});
        off += dbf[o + 16];
    }
    if (field_index >= (int)fields.size()) {
        err = "field index " + std::to_string(field_index) + " out of range (" + std::to_string(fields.size()) +
              " fields)";
        return false;
    }
    if ((size_t)hlen + (size_t)nrec * rlen > dbf.size()) {
        err = "truncated DBF file";
        return false;
    }
    if (field_index >= 0 && fields[field_index].off + fields[field_index].len > (int)rlen) {
        err = "DBF field " + std::to_string(field_index) + " lies outside the record";
        return false;
    }
    // walk the .shp records alongside the .dbf records
    pargo::alloc_point("shapefile.points");
    out.x.reserve(nrec);
    out.y.reserve(nrec);
    out.value.reserve(nrec);
    size_t o = 100;
    for (uint32_t rec = 0; rec < nrec; rec++) {
        if (o + 8 > shp.size()) {
            err = "truncated SHP file";
            return false;
        }
        const int64_t clen = (int64_t)be32(&shp[o + 4]) * 2;  // 16-bit words
        if (clen < 4 || (uint64_t)clen > shp.size() - (o + 8)) {
            err = "truncated SHP record";
            return false;
        }
        const unsigned char *c = &shp[o + 8];
        o += 8 + (size_t)clen;
        const unsigned char *r = &dbf[hlen + (size_t)rec * rlen];
        if (r[0] == '*') continue;  // deleted record
        const int32_t st = le32(c);
        if (!(st == 1 || st == 11 || st == 21) || clen < 20) {
            err = "[ERROR] No point geometry";
            return false;
        }
        out.x.push_back(led(c + 4));
        out.y.push_back(led(c + 12));
        if (field_index < 0) {
            out.value.push_back(0.0);
            continue;
        }
        const Field &f = fields[field_index];
        std::string txt((const char *)r + f.off, (size_t)f.len);
        out.value.push_back(std::strtod(txt.c_str(), nullptr));
    }
    return true;
}

}  // namespace shapefile
}  // namespace GPRO
