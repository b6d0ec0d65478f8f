// Fitted walls: the cut of a closed triangle mesh by the grid's cells, and each air cell's
// share of it (its air volume, the open area towards each of its 18 neighbours, its wall
// area per material, the slivers of air it owns in its face neighbours' cells). It stands
// in for the reference's CutCells (cpp/include/acoustic_spaces/cut_cells.hpp) on meshes
// whose faces close (every edge met by an even count of faces inside the grid), lie off the
// grid's planes and off its cells' corners; see fitted.cpp for the design.

#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace fitted {

constexpr uint8_t kSolid = 0, kAir = 1, kLayer = 2;

struct Mesh {
    const double *vertices = nullptr;  // vertex_count x 3
    size_t vertex_count = 0;
    const int64_t *faces = nullptr;    // face_count x 3
    size_t face_count = 0;
};

struct GridSpec {
    double origin[3] = {0, 0, 0};
    double cell = 0;
    size_t n[3] = {0, 0, 0};
    size_t layer = 0;
};

// A point on the lattice and its bracket among the cells' centres along each axis: j the
// last centre not above it (-1 below the first), tied where it is that centre, lo and hi its
// fractions of the way from centre j to centre j + 1 and back, each correctly rounded (0 and
// 1 where j has no centre above).
struct Located {
    int64_t q[3] = {0, 0, 0};
    long long j[3] = {0, 0, 0};
    bool tied[3] = {false, false, false};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
};

// Counts of the fit, for tests and the design notes.
struct FitStats {
    size_t faces = 0, planes = 0, cubes = 0, leaves = 0, regions = 0, squares = 0, ties = 0;
    size_t farthest = 0, conduit_links = 0, conduit_scaled = 0;
    double far_volume = 0, diagonal_area = 0, dropped_area = 0, edge_area = 0, air_volume = 0;
    double conduit_least = 1.0, conduit_wall_moved = 0;
    int quantum_bits = 0;
};

class Fit {
public:
    // Throws std::invalid_argument with a message where the input is refused.
    Fit(const Mesh &mesh, const int64_t *face_material, size_t materials, const GridSpec &grid,
        const double source[3], int threads);
    ~Fit();

    // Each cell's kind (kSolid, kAir, kLayer), x fastest.
    std::vector<uint8_t> kind;
    // The rows, ascending by cell: air volume (cells^3), open area per direction (n x 18,
    // cell faces), wall area per material then the grid's edge (n x (materials + 1)),
    // sliver thickness through each axial face (n x 6, cells).
    std::vector<int64_t> rows;
    std::vector<double> air_volume, open_area, wall_area, sliver;
    // Each cell's width along each axis on the lattice, in cells.
    std::array<std::vector<double>, 3> width;
    std::array<bool, 6> open_sides{};
    FitStats stats;

    double cell_volume(size_t i) const;
    double face_area(size_t i, int axis) const;

    Located locate(const double p[3]) const;
    void centre_q(size_t i, int64_t out[3]) const;
    // Whether lattice point q lies in the grid's closed box.
    bool in_grid_box(const int64_t q[3]) const;
    // The first face (the mesh's index) the segment from a to b crosses, or -1. Throws
    // std::invalid_argument where it passes exactly through a face's edge or an end lies
    // on a face's plane inside the face.
    int64_t way(const int64_t a[3], const int64_t b[3]) const;
    // Whether lattice point q lies on a face.
    bool on_face(const int64_t q[3]) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace fitted
