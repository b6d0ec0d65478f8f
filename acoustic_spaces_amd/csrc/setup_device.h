// The set-up's passes on the device (setup_device.hip), as the host set-up (setup.cpp)
// calls them: plain structs, no HIP types.
#ifndef AS_SETUP_DEVICE_H
#define AS_SETUP_DEVICE_H

#include <cstdint>
#include <vector>

namespace asdev {

// A mesh face for the crossing test: a vertex, two edges from it, the unit normal, its box.
struct Tri {
    double v0[3], e1[3], e2[3], n[3], lo[3], hi[3];
};

// The grid, the triangle bins' shape, the margins and the stencil's directions.
struct Grid {
    double origin[3], cell, guard, barycentric;
    uint32_t nx, ny, nz, bin, bx, by, bz;
    uint32_t directions;  // 6 or 18
    int8_t step[18][3];
    int8_t plus_bit[18];  // the direction's bit in a cell's cut word, or -1
    int8_t opposite[18];
};

// On the current device: every cell's cut word (cut, n values), then, unless a flag is
// raised, the air (air: n bytes, 1 at the seeds, on return 1 at every cell they reach through
// open links; rounds: the flood's rounds; seconds, where not null: the two passes' times). The flags: bit 0, an end lies within the guard of a
// face's plane next to the face; bit 1, a link meets a face only within the margins and no
// face cuts it firmly. Throws std::invalid_argument on inconsistent arguments and
// std::runtime_error on a HIP failure.
uint32_t cut_and_flood(const std::vector<Tri> &tri, const std::vector<uint32_t> &bin_start,
                       const std::vector<uint32_t> &bin_faces, const Grid &g, std::vector<uint16_t> &cut,
                       std::vector<uint8_t> &air, size_t &rounds, double *seconds);

}  // namespace asdev

#endif
