// Minimal FITS reader for HEALPix maps (binary-table extension).
//
// FITS layout: 2880-byte records; a header is a run of 80-character cards
// "KEYWORD = value / comment" closed by an END card and padded to a record;
// the data follow, big-endian, padded to a record.  The primary HDU of a
// HEALPix file normally has no data (NAXIS = 0).
#include "healpix_fits.h"

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace opsec_host {

namespace {

constexpr size_t RECORD = 2880;

struct Header {
    std::map<std::string, std::string> keys;   // values with quotes and comments removed
    long long get_int(const std::string& k, long long def = -1) const {
        auto it = keys.find(k);
        return it == keys.end() ? def : std::strtoll(it->second.c_str(), nullptr, 10);
    }
    double get_double(const std::string& k, double def) const {
        auto it = keys.find(k);
        if (it == keys.end()) return def;
        std::string v = it->second;
        for (char& c : v)
            if (c == 'D') c = 'E';   // FITS allows D exponents
        return std::strtod(v.c_str(), nullptr);
    }
    bool has(const std::string& k) const { return keys.count(k) != 0; }
    std::string get(const std::string& k) const {
        auto it = keys.find(k);
        return it == keys.end() ? std::string() : it->second;
    }
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Reads one header (whole records) from f; false at end of file or on a
// header without END.
bool read_header(FILE* f, Header* h) {
    char rec[RECORD];
    for (;;) {
        if (std::fread(rec, 1, RECORD, f) != RECORD) return false;
        for (size_t c = 0; c < RECORD; c += 80) {
            std::string card(rec + c, 80);
            std::string key = trim(card.substr(0, 8));
            if (key == "END") return true;
            if (card.size() < 10 || card[8] != '=' || card[9] != ' ') continue;   // COMMENT, HISTORY, blank
            std::string val = card.substr(10);
            size_t p = val.find_first_not_of(' ');
            if (p != std::string::npos && val[p] == '\'') {
                // quoted string; '' is an escaped quote
                std::string s;
                size_t i = p + 1;
                for (; i < val.size(); i++) {
                    if (val[i] == '\'') {
                        if (i + 1 < val.size() && val[i + 1] == '\'') {
                            s += '\'';
                            i++;
                        } else {
                            break;
                        }
                    } else {
                        s += val[i];
                    }
                }
                h->keys[key] = trim(s);
            } else {
                size_t slash = val.find('/');
                h->keys[key] = trim(slash == std::string::npos ? val : val.substr(0, slash));
            }
        }
    }
}

size_t padded(size_t n) { return (n + RECORD - 1) / RECORD * RECORD; }

// Bytes of the data unit of an HDU (FITS standard 4.4.1).
size_t data_bytes(const Header& h) {
    const long long naxis = h.get_int("NAXIS", 0);
    if (naxis <= 0) return 0;
    long long n = 1;
    for (long long i = 1; i <= naxis; i++) n *= h.get_int("NAXIS" + std::to_string(i), 0);
    const long long bitpix = std::llabs(h.get_int("BITPIX", 8));
    const long long gcount = h.get_int("GCOUNT", 1), pcount = h.get_int("PCOUNT", 0);
    return (size_t)(bitpix / 8 * gcount * (pcount + n));
}

uint64_t be_read(const unsigned char* p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 8) | p[i];
    return v;
}

}  // namespace

bool read_healpix_map(const std::string& path, HealpixMap* out, std::string* why) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        *why = "could not open '" + path + "'";
        return false;
    }
    auto fail = [&](const std::string& m) {
        std::fclose(f);
        *why = m + " in '" + path + "'";
        return false;
    };
    Header prim, ext;
    if (!read_header(f, &prim) || prim.get("SIMPLE") != "T") return fail("not a FITS file");
    if (std::fseek(f, (long)padded(data_bytes(prim)), SEEK_CUR) != 0) return fail("truncated primary HDU");
    if (!read_header(f, &ext)) return fail("no extension HDU");
    if (ext.get("XTENSION") != "BINTABLE") return fail("extension is not binary");
    const long long rowbytes = ext.get_int("NAXIS1"), nrows = ext.get_int("NAXIS2");
    const long long nside = ext.get_int("NSIDE", 0);
    if (rowbytes <= 0 || nrows <= 0) return fail("bad table size");
    // nside 16384 is 3.2e9 pixels (12.9 GB as float); larger maps are refused
    if (nside <= 0 || nside > (1LL << 14)) return fail("missing or bad NSIDE");
    const long long npix = 12 * nside * nside;
    if (npix % nrows != 0) return fail("Problem with npix");
    const long long per_row = npix / nrows;
    // column 1: repeat count and type, at byte 0 of each row
    std::string tform = ext.get("TFORM1");
    size_t t = 0;
    while (t < tform.size() && std::isdigit((unsigned char)tform[t])) t++;
    const long long repeat = t ? std::strtoll(tform.substr(0, t).c_str(), nullptr, 10) : 1;
    const char type = t < tform.size() ? tform[t] : '?';
    int width;
    switch (type) {
        case 'E': width = 4; break;
        case 'D': width = 8; break;
        case 'J': width = 4; break;
        case 'I': width = 2; break;
        case 'K': width = 8; break;
        case 'B': width = 1; break;
        default: return fail("unsupported column type '" + tform + "'");
    }
    if (repeat < per_row || repeat * width > rowbytes) return fail("column 1 holds fewer values than npix / NAXIS2");
    const double scal = ext.get_double("TSCAL1", 1.0), zero = ext.get_double("TZERO1", 0.0);
    // the table must be in the file before anything is allocated for it
    const long here = std::ftell(f);
    if (std::fseek(f, 0, SEEK_END) != 0) return fail("unreadable file");
    const long end = std::ftell(f);
    if (here < 0 || end < here || (unsigned long long)(end - here) < (unsigned long long)rowbytes * nrows)
        return fail("truncated table");
    if (std::fseek(f, here, SEEK_SET) != 0) return fail("unreadable file");
    std::vector<unsigned char> row((size_t)rowbytes);
    out->map.assign((size_t)npix, 0.0f);
    for (long long r = 0; r < nrows; r++) {
        if (std::fread(row.data(), 1, row.size(), f) != row.size()) return fail("truncated table");
        for (long long i = 0; i < per_row; i++) {
            const unsigned char* p = row.data() + i * width;
            double v;
            const uint64_t b = be_read(p, width);
            if (type == 'E') {
                uint32_t u = (uint32_t)b;
                float x;
                std::memcpy(&x, &u, 4);
                v = x;
            } else if (type == 'D') {
                double x;
                std::memcpy(&x, &b, 8);
                v = x;
            } else if (type == 'J') {
                v = (double)(int32_t)(uint32_t)b;
            } else if (type == 'I') {
                v = (double)(int16_t)(uint16_t)b;
            } else if (type == 'K') {
                v = (double)(int64_t)b;
            } else {
                v = (double)(uint8_t)b;
            }
            v = zero + scal * v;
            out->map[(size_t)(r * per_row + i)] = std::isnan(v) ? -1.6375e30f : (float)v;
        }
    }
    std::fclose(f);
    out->nside = nside;
    out->ordering = ext.get("ORDERING");
    out->coordsys = ext.get("COORDSYS");
    return true;
}

}  // namespace opsec_host
