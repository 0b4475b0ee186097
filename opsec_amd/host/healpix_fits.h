// HEALPix completeness maps stored as FITS binary tables, the maskfile of a
// SphericalSurvey (read_healpix_map, src/chealpix.c:411-480, as used by
// src/SphericalSurvey.cpp:88-124).
#pragma once

#include <string>
#include <vector>

namespace opsec_host {

struct HealpixMap {
    long long nside = 0;
    std::string ordering;   // "RING" or "NESTED" (ORDERING keyword)
    std::string coordsys;   // COORDSYS keyword, if present
    std::vector<float> map; // 12 nside^2 values
};

// The map in the first extension (HDU 2), which must be a binary table: NSIDE
// from its header, 12 NSIDE^2 values read row by row from column 1 (NAXIS2
// rows of npix / NAXIS2 values each), converted to float; NaN values become
// the HEALPix null value -1.6375e30.  Column types E, D, J, I, K and B are
// read, with TSCAL1 / TZERO1 applied; NSIDE up to 16384.  Returns false with
// *why set on any problem (checked before the map is allocated).
bool read_healpix_map(const std::string& path, HealpixMap* out, std::string* why);

}  // namespace opsec_host
