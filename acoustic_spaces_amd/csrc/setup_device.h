// The set-up's passes over the whole grid on the device (setup_device.hip), as the host
// set-up (setup.cpp) calls them: plain structs and vectors, no HIP types.
#ifndef AS_SETUP_DEVICE_H
#define AS_SETUP_DEVICE_H

#include <cstdint>
#include <vector>

namespace asdev {

// A mesh face: a vertex, two edges from it, the unit normal, its box, its material and the
// bins it is listed in (from, to along each axis, inclusive).
struct Tri {
    double v0[3], e1[3], e2[3], n[3], lo[3], hi[3];
    int64_t material;
    uint32_t bin_from[3], bin_to[3];
};

// The grid, the triangle bins' shape, the margins, the stencil's directions and the
// constants of the wall cells.
struct Grid {
    double origin[3], cell, guard, barycentric;
    uint32_t nx, ny, nz, bin, bx, by, bz;
    uint32_t directions;  // 6 or 18
    uint32_t layer;       // the matched layer's width in cells (0: none)
    double edge_admittance;
    double exact[18];     // each direction's weight
    double wall[18];      // and its boundary weight
    int8_t step[18][3];
    int8_t plus_bit[18];  // the direction's bit in a cell's cut word, or -1
    int8_t opposite[18];
    int8_t side_a[18], side_b[18];  // a diagonal's axial components, lower axis first
};

// What the set-up hands the device.
struct Input {
    Grid g;
    std::vector<Tri> tri;
    std::vector<uint32_t> bin_start, bin_faces;
    std::vector<uint8_t> seeds;           // n bytes: 1 at the cells the source sees
    const float *scale = nullptr;         // n values (a cell's speed squared), or null
    std::vector<double> beta;             // per material
    std::vector<uint32_t> branch_start;   // per material, into branch_of (materials + 1)
    std::vector<int32_t> branch_of;       // the branches of each material, in order
    std::vector<double> branch_a;         // per branch: m + r / 2 + q / 2
};

// Where a pass refused the geometry. The wall cells' refusal is the one at the lowest cell.
enum Fault : uint32_t {
    kFine = 0,
    kGuard = 1,           // an end within the guard of a face's plane next to the face
    kEdgeLink = 2,        // a link meets a face only within the margins, none cuts it firmly
    kLayerReach = 3,      // a face cuts a link between two cells of the matched layer
    kManyEdges = 4,       // a segment passes through more than 16 faces' edges
    kTwice = 5,           // a segment crosses two faces of different materials at one point
    kEdgeOnly = 6,        // a segment passes through a face's edge and crosses none
    kEdgeNear = 7,        // a segment passes through a face's edge next to its nearest crossing
};

// What the device hands back: the cut words, each cell's class, the layer's cells and the
// wall cells' arrays, as the host set-up builds them.
struct Output {
    uint32_t fault = kFine;
    uint64_t fault_cell = 0;
    size_t rounds = 0;
    double seconds[4] = {0, 0, 0, 0};  // cut links, flood, layer and marks, wall cells
    std::vector<uint16_t> cut;
    std::vector<uint8_t> type;
    std::vector<int64_t> layer_cells;
    uint32_t sides = 0;                 // bit k: side k (-x, +x, -y, +y, -z, +z) open
    bool open = false;
    uint64_t interior = 0;              // interior cells before the wall cells are taken out
    std::vector<int64_t> cell;
    std::vector<uint64_t> link_start, state_start;
    std::vector<uint8_t> link_direction;
    std::vector<float> link_weight, coef_a, coef_b, coef_inv, state_weight;
    std::vector<int32_t> state_branch;
};

// On the current device: the cut links, the air's flood from the seeds, the matched layer,
// the wall cells' marks and the wall cells. Stops at the first pass that refuses (fault set;
// the arrays then hold what that pass left). Throws std::invalid_argument on inconsistent
// arguments and std::runtime_error on a HIP failure.
void set_up(const Input &in, Output &out);

}  // namespace asdev

#endif
