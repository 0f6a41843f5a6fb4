// geometry.hpp -- plasma-region geometry of the implicit NDD step: the
// reference's block boxes and boundary segments turned into per-face type
// codes, ghost rules and electrode faces (host side).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "cathode_hip.h"

namespace cathode {

// face type codes (x faces: low = the face of the plasma cell on its low-x
// side, i.e. plasma at i, wall at face i; high = plasma at i - 1; r faces
// likewise); GF_BIAS marks a wall at V_bias
enum : uint8_t {
    GF_NONE = 0,
    GF_INNER = 1,
    GF_WLO = 2,
    GF_WHI = 3,
    GF_OPEN = 4,
    GF_AXIS = 5,
    GF_TYPE = 0x0f,
    GF_BIAS = 0x10,
    GF_INLET = 0x20  // a high-r wall face of the neutral-gas inlet segment
};

// per-cell ghost writes after a routine (bit d: write the neighbour on side
// d = 0 low x, 1 high x, 2 low r, 3 high r; bit 4 + d: copy the cell, else 0)
enum : uint8_t { GW_XLO = 1, GW_XHI = 2, GW_RLO = 4, GW_RHI = 8 };

struct GeoTables {
    int ni = 0, nj = 0, ld = 0;         // ld = nj + 2
    std::vector<uint8_t> xc, rc, act;   // x-face codes, r-face codes, plasma cells
    // potential ghost rule per cell of the extended grid: -1, or
    // (source cell << 2) | (Dirichlet << 1) | bias, the last segment's
    std::vector<int32_t> gs;
    std::vector<uint8_t> ne, te;        // ghost writes of n_e and Te per plasma cell
    std::vector<uint8_t> blk;           // 1 + the block index of each plasma cell, 0 elsewhere
    std::vector<int> anode;             // biased high-x walls: (face i, j0, j1) triples
    int ncells = 0;                     // plasma cells
};

// the rectangle 1..ni x 1..nj with the block-4 boundary set
cathode_ndd_geometry rect_geometry(int ni, int nj);

// validates g on the ni x nj grid and fills T; false with the reason in err
bool build_geo_tables(int ni, int nj, const cathode_ndd_geometry& g, GeoTables& T, std::string& err);

}  // namespace cathode
