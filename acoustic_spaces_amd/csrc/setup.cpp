// The wave solver's set-up on walls on cell faces: from a triangle mesh, its materials and
// a grid, the layout the time step reads. It stands in for the reference's WaveSolver
// constructor on that path (cpp/include/acoustic_spaces/fdtd.hpp:360-690) and hands out
// the arrays of its layout() (727-782), equal to the reference's bit for bit.
//
// How it is organised:
//  - Stencil: the 18 link directions as a table of their steps, weights and boundary
//    weights; a diagonal's two axial sides and each direction's opposite are derived from
//    the steps.
//  - The geometry: triangles binned in blocks of cells; a segment crosses a face when its
//    ends lie strictly on either side of the face's plane and the point where it meets
//    the plane lies inside the triangle.
//  - The cut word of each cell (a bit per link of the + half: the direction whose cell
//    offset is positive owns the bit) and, from it, each cell's LinkMask: which of its
//    links are blocked (cut, or reaching a cell on the grid's edge).
//  - The air: a span flood from the cells the source sees.
//  - Wall cells from LinkMasks alone: the folded weights of a cell are a function of its
//    mask; the marks, the pair-averaged weights, the wall contacts of each blocked link,
//    their resolution to a material or the grid's edge, and the sum of wall term and
//    branch states are separate stages, run on slabs of planes on all threads.
//  - SpecialList, which builds the layout's arrays of special cells.
//  - Readings: the source's and probes' stencils over a 4 x 4 x 4 block as 64-bit masks.
//
// What parity fixes, and is therefore the reference's: the decisions' rules (which links
// are blocked, which cells are marked, which wall a contact meets), every formula and its
// operation order (the lowered and rounded-down weights, fdtd.hpp:2127-2135; the
// pair-averaged weights; the wall term's sum in contact order; the branch coefficients,
// 391-411; the specials' coefficients, 3499-3517; the Lagrange and trilinear weights,
// 2361-2488 and 3305-3365; the matched layer's profile, 2219-2240) and the order of the
// output arrays (specials by ascending cell, then the source's cells; taps in z, y, x
// order).
//
// Geometry and ties. The reference decides crossings with a BVH and tie rules for points
// on faces and segments through their edges (kTie, crossing, fdtd.hpp:2475-3257). Here
// every decision is made with a margin (kGuard from a plane, kBarycentric from an edge)
// far wider than the reference's tie and edge tolerances (1e-10 cells, 1e-9 barycentric);
// a decision inside a margin that could change the layout raises "degenerate" instead of
// guessing, so every layout this code hands out is the reference's. A meeting near an edge
// shared with another face of the same material at the same point, or farther than a firm
// crossing, changes nothing and is accepted. Sources and probes on faces (the reference's
// leans) are refused.

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "asfdtd.h"
#include "fitted.h"
#include "setup_device.h"

namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
    double operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
inline V3 operator+(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3 &a, const V3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3 &a) { return std::sqrt(dot(a, a)); }

// The set-up's phases' times on stderr, where AS_FDTD_TIMING is set.
struct SetupClock {
    bool on = std::getenv("AS_FDTD_TIMING") != nullptr;
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void mark(const char *what) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "as_setup_create: %-22s %9.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    }
};

struct Degenerate : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

constexpr uint8_t kSolid = 0, kInterior = 1, kSpecial = 2, kLayer = 3;
constexpr double kMargin = 1.0 / 262144.0;  // the links' stability margin, 2^-18 (fdtd.hpp:17-27)
constexpr int kReach = 4;                   // cells a side of a cubic reading (fdtd.hpp:2321)
constexpr double kNear = 3.0;               // probes nearer the source (in cells) read trilinearly
constexpr double kBarycentric = 1e-8;       // the edge margin (see above)
constexpr size_t npos = static_cast<size_t>(-1);

// The link directions, the layout's link_direction values: 6 axial, then 12 diagonal.
constexpr int kSteps[18][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

// The largest float not above w, and a link's weight lowered by the margin then rounded
// so (fdtd.hpp:2127-2135): no weight rounds up, so the links never sum above their bound.
float down(double w) {
    float f = static_cast<float>(w);
    if (static_cast<double>(f) > w) f = std::nextafter(f, 0.0f);
    return f;
}
float lowered(double w) { return down(w * (1.0 - kMargin)); }

// --- the stencil ----------------------------------------------------------------------

struct Direction {
    int step[3] = {0, 0, 0};
    ptrdiff_t offset = 0;  // in cells of the flat grid
    double exact = 0;      // the link's weight (0 for a direction the stencil does not use)
    float weight = 0;      // lowered to float
    double wall = 0;       // its boundary weight: exact / (2 L), L the Courant number
    int plus_bit = -1;     // its bit in a cell's cut word, where its offset is positive
    int opposite = 0;
    int side_a = -1, side_b = -1;  // a diagonal's axial components, lower axis first
};

// The 7-point stencil (axial 1/3, L = 1/sqrt 3) or the 19-point one (axial 1/4, diagonal
// 1/8, L = sqrt(3/4)) (fdtd.hpp:7-15, set_up_stencil 1649-1673).
struct Stencil {
    int count = 6;  // directions used: 6 or 18
    std::array<Direction, 18> d;
    float centre = 0, total = 0;  // 2 less the weights' sum, and the sum, as floats
};

Stencil make_stencil(size_t points, ptrdiff_t sy, ptrdiff_t sz) {
    Stencil st;
    st.count = points == 7 ? 6 : 18;
    const double axial = points == 7 ? 1.0 / 3.0 : 0.25;
    const double diagonal = points == 7 ? 0.0 : 0.125;
    const double courant = std::sqrt(axial + 4.0 * diagonal);
    int bits = 0;
    for (int k = 0; k < 18; ++k) {
        Direction &d = st.d[static_cast<size_t>(k)];
        int span = 0;
        for (int a = 0; a < 3; ++a) {
            d.step[a] = kSteps[k][a];
            span += std::abs(d.step[a]);
        }
        d.offset = d.step[0] + d.step[1] * sy + d.step[2] * sz;
        d.exact = k < st.count ? (span == 1 ? axial : diagonal) : 0.0;
        d.weight = lowered(d.exact);
        d.wall = d.exact / (2.0 * courant);
        if (k < st.count && d.offset > 0) d.plus_bit = bits++;
    }
    for (int k = 0; k < 18; ++k) {
        Direction &d = st.d[static_cast<size_t>(k)];
        for (int j = 0; j < 18; ++j) {
            const int *s = st.d[static_cast<size_t>(j)].step;
            if (s[0] == -d.step[0] && s[1] == -d.step[1] && s[2] == -d.step[2]) d.opposite = j;
        }
        if (k < 6) continue;
        for (int a = 0; a < 3; ++a) {
            if (d.step[a] == 0) continue;
            const int axial_dir = 2 * a + (d.step[a] > 0 ? 1 : 0);
            (d.side_a < 0 ? d.side_a : d.side_b) = axial_dir;
        }
    }
    double sum = 0.0;
    for (int k = 0; k < st.count; ++k) sum += st.d[static_cast<size_t>(k)].weight;
    st.centre = static_cast<float>(2.0 - sum);
    st.total = static_cast<float>(sum);
    return st;
}

// Which of a cell's links are blocked: bit k where its neighbour along k lies on the grid's
// edge or a face cuts the link; edge_far, the first kind alone.
struct LinkMask {
    uint32_t blocked = 0, edge_far = 0;
};

// A cell's link weights before they are averaged with its neighbours' (raw_weights,
// fdtd.hpp:2162-2190), from its mask alone: an open link keeps its weight; a blocked
// diagonal whose far cell is not on the grid's edge hands its weight to the open axial
// links it is made of, shared equally; every other blocked link weighs 0. (The shares are
// sums of binary fractions, exact in any order.)
void fold(const Stencil &st, const LinkMask &m, double *w) {
    for (int k = 0; k < 18; ++k) {
        const bool open = k < st.count && !((m.blocked >> k) & 1u);
        w[k] = open ? st.d[static_cast<size_t>(k)].exact : 0.0;
    }
    for (int k = 6; k < st.count; ++k) {
        if (!((m.blocked >> k) & 1u) || ((m.edge_far >> k) & 1u)) continue;
        const Direction &d = st.d[static_cast<size_t>(k)];
        const bool open_a = !((m.blocked >> d.side_a) & 1u), open_b = !((m.blocked >> d.side_b) & 1u);
        const int sides = (open_a ? 1 : 0) + (open_b ? 1 : 0);
        if (sides == 0) continue;
        const double share = d.exact / sides;
        if (open_a) w[d.side_a] += share;
        if (open_b) w[d.side_b] += share;
    }
}

// --- the matched layer's profile ----------------------------------------------------------

// Its coefficients along one axis of `cells` cells (layer_coefficients, fdtd.hpp:2219-2240),
// at x = 0 .. cells: u the distance in cells from the nearer end, of the centre
// (min(x, last - x)) and of the lower face (min(x - 1/2, last - x + 1/2)); the damping
// d(u) = damping ((width + 1/2 - u) / width)^2 where that is positive, else 0; keep =
// (1 - d) / (1 + d), push = courant / (1 + d).
struct LayerAxis {
    std::vector<float> keep_centre, keep_face, push_centre, push_face;
};

LayerAxis layer_axis(size_t cells, size_t width, double damping, double courant) {
    LayerAxis out;
    const double w = static_cast<double>(width), last = static_cast<double>(cells - 1);
    const auto profile = [&](double u) {
        const double depth = std::max(0.0, w + 0.5 - u) / w;
        return damping * depth * depth;
    };
    for (size_t x = 0; x <= cells; ++x) {
        const double c = static_cast<double>(x);
        const double at_centre = profile(std::min(c, last - c));
        const double at_face = profile(std::min(c - 0.5, last - c + 0.5));
        out.keep_centre.push_back(static_cast<float>((1.0 - at_centre) / (1.0 + at_centre)));
        out.push_centre.push_back(static_cast<float>(courant / (1.0 + at_centre)));
        out.keep_face.push_back(static_cast<float>((1.0 - at_face) / (1.0 + at_face)));
        out.push_face.push_back(static_cast<float>(courant / (1.0 + at_face)));
    }
    return out;
}

// --- branches ---------------------------------------------------------------------------

// Each admittance branch (m, r, q) centred in time (fdtd.hpp:391-411): a = m + r/2 + q/2,
// now = 2m / a, before = (2m + q) / a, gain = 1 / (2a); and each material's branches.
struct Branches {
    std::vector<float> now, before, gain;
    std::vector<double> a;
    std::vector<std::vector<int32_t>> of_material;
};

Branches make_branches(const double *mrq, const int64_t *material, size_t count, size_t materials) {
    Branches b;
    b.of_material.resize(materials);
    for (size_t j = 0; j < count; ++j) {
        const double m = mrq[3 * j], r = mrq[3 * j + 1], q = mrq[3 * j + 2];
        const double a = m + 0.5 * r + 0.5 * q;
        if (!(m >= 0.0 && r >= 0.0 && q >= 0.0 && a > 0.0)) {
            throw std::invalid_argument("branch " + std::to_string(j) +
                                        ": mass, resistance and stiffness must be at least 0, and not all 0");
        }
        if (material[j] < 0 || static_cast<size_t>(material[j]) >= materials) {
            throw std::invalid_argument("branch " + std::to_string(j) + " belongs to no material");
        }
        b.of_material[static_cast<size_t>(material[j])].push_back(static_cast<int32_t>(j));
        b.now.push_back(static_cast<float>(2.0 * m / a));
        b.before.push_back(static_cast<float>((2.0 * m + q) / a));
        b.gain.push_back(static_cast<float>(0.5 / a));
        b.a.push_back(a);
    }
    return b;
}

// --- the special cells' arrays -----------------------------------------------------------

// The layout's arrays of special cells (fdtd.hpp:727-782), appended a cell at a time:
// its links where its weight is positive, lowered; coef_a = 2 - s (sum of the weights)
// (1 - 2^-18); coef_b = 1 - r g and coef_inv = 1 / (1 + r g), r = sqrt s, g its wall term;
// its branch states (add_special, fdtd.hpp:3499-3517).
class SpecialList {
public:
    std::vector<int64_t> cell;
    std::vector<uint64_t> link_start{0};
    std::vector<uint8_t> link_direction;
    std::vector<float> link_weight;
    std::vector<float> coef_a, coef_b, coef_inv, source;
    std::vector<uint64_t> state_start{0};
    std::vector<int32_t> state_branch;
    std::vector<float> state_weight;
    size_t sorted = 0;  // the first `sorted` cells ascend (the wall cells)

    struct State {
        int32_t branch;
        float weight;
    };

    void append(size_t i, const double *w, int count, double s, double g, const State *states, size_t nstates) {
        cell.push_back(static_cast<int64_t>(i));
        double sum = 0.0;
        for (int k = 0; k < count; ++k) {
            if (!(w[k] > 0.0)) continue;
            sum += w[k];
            link_direction.push_back(static_cast<uint8_t>(k));
            link_weight.push_back(lowered(s * w[k]));
        }
        link_start.push_back(link_direction.size());
        const double r = std::sqrt(s);
        coef_a.push_back(static_cast<float>(2.0 - s * sum * (1.0 - kMargin)));
        coef_b.push_back(static_cast<float>(1.0 - r * g));
        coef_inv.push_back(static_cast<float>(1.0 / (1.0 + r * g)));
        source.push_back(0.0f);
        for (size_t q = 0; q < nstates; ++q) {
            state_branch.push_back(states[q].branch);
            state_weight.push_back(states[q].weight);
        }
        state_start.push_back(state_branch.size());
    }

    // Appends a cell whose stored weights and coefficients are given (fitted walls): a link
    // for each direction whose area is positive.
    void append_given(size_t i, const double *area, const float *weight, float a, float b, float inv,
                      const State *states, size_t nstates) {
        cell.push_back(static_cast<int64_t>(i));
        for (int k = 0; k < 18; ++k) {
            if (!(area[k] > 0.0)) continue;
            link_direction.push_back(static_cast<uint8_t>(k));
            link_weight.push_back(weight[k]);
        }
        link_start.push_back(link_direction.size());
        coef_a.push_back(a);
        coef_b.push_back(b);
        coef_inv.push_back(inv);
        source.push_back(0.0f);
        for (size_t q = 0; q < nstates; ++q) {
            state_branch.push_back(states[q].branch);
            state_weight.push_back(states[q].weight);
        }
        state_start.push_back(state_branch.size());
    }

    // Appends another list's cells (their ranges shifted by this list's counts).
    void extend(const SpecialList &o) {
        const uint64_t links0 = link_direction.size(), states0 = state_branch.size();
        cell.insert(cell.end(), o.cell.begin(), o.cell.end());
        for (size_t q = 1; q < o.link_start.size(); ++q) link_start.push_back(links0 + o.link_start[q]);
        for (size_t q = 1; q < o.state_start.size(); ++q) state_start.push_back(states0 + o.state_start[q]);
        link_direction.insert(link_direction.end(), o.link_direction.begin(), o.link_direction.end());
        link_weight.insert(link_weight.end(), o.link_weight.begin(), o.link_weight.end());
        for (auto [to, from] : {std::pair{&coef_a, &o.coef_a}, std::pair{&coef_b, &o.coef_b},
                                std::pair{&coef_inv, &o.coef_inv}, std::pair{&source, &o.source},
                                std::pair{&state_weight, &o.state_weight}})
            to->insert(to->end(), from->begin(), from->end());
        state_branch.insert(state_branch.end(), o.state_branch.begin(), o.state_branch.end());
    }

    // The special index of cell i, or npos: a binary search of the ascending wall cells,
    // then the few cells added after them.
    size_t index_of(size_t i) const {
        const auto key = static_cast<int64_t>(i);
        const auto end = cell.begin() + static_cast<ptrdiff_t>(sorted);
        const auto it = std::lower_bound(cell.begin(), end, key);
        if (it != end && *it == key) return static_cast<size_t>(it - cell.begin());
        for (size_t k = sorted; k < cell.size(); ++k)
            if (cell[k] == key) return k;
        return npos;
    }

    size_t size() const { return cell.size(); }
};

// --- readings ----------------------------------------------------------------------------

// The weight of node j of four (0 .. 3) in the cubic Lagrange interpolation at x, the
// product over the other nodes m, in ascending m, of (x - m) / (j - m) (fdtd.hpp:2361-2488).
double lagrange4(double x, int j) {
    double w = 1.0;
    for (int m = 0; m < kReach; ++m)
        if (m != j) w *= (x - m) / static_cast<double>(j - m);
    return w;
}

// A 4 x 4 x 4 block of cells as a 64-bit set, bit x + 4 y + 16 z.
constexpr uint64_t block_bit(int x, int y, int z) { return 1ull << (x + 4 * y + 16 * z); }

// The cells of the block whose coordinate along axis a is in `planes` (a 4-bit set).
uint64_t block_slab(int a, unsigned planes) {
    uint64_t out = 0;
    for (int z = 0; z < kReach; ++z)
        for (int y = 0; y < kReach; ++y)
            for (int x = 0; x < kReach; ++x) {
                const int at[3] = {x, y, z};
                if ((planes >> at[a]) & 1u) out |= block_bit(x, y, z);
            }
    return out;
}

// --- triangles ---------------------------------------------------------------------------

struct Triangle {
    V3 v0, e1, e2, normal;
    V3 lo, hi;  // bounding box
    int64_t material = 0;
};

// A segment's meeting with the faces: the nearest firm crossing.
struct Hit {
    double t = -1.0;
    int64_t face = -1;
};

// What a wall contact meets: a material's face (>= 0), the grid's edge, or nothing (a link
// that only grazes a face's edge counts as rigid, fdtd.hpp:592-593).
constexpr int32_t kEdgeWall = -1, kNoWall = -2;

class Setup {
public:
    // inputs
    std::vector<Triangle> tri;
    std::vector<int64_t> face_material;
    std::vector<double> beta;
    size_t materials = 0;
    V3 origin;
    double cell = 0;
    size_t nx = 0, ny = 0, nz = 0, sy = 0, sz = 0, n = 0;
    size_t layer = 0;
    double layer_damping = 0;
    double edge_admittance = 1.0;
    size_t points = 7;
    int threads = 1;
    double guard = 0;  // kGuard: distance from a plane within which an end counts as on it
    bool on_device = false;  // the cut links and the flood on the current GPU
    size_t flood_rounds = 0;  // the device flood's rounds
    // Moving air (see set_up_flow): the shifts a step along x and y and each cell's row
    // mask (plain), or the wind on each cell's +x, +y, +z faces and the loops (round); the
    // mask then 1 where the air moves the cell, 2 at the border of the moving air.
    std::vector<float> flow_x, flow_y;
    std::vector<uint8_t> flow_mask;
    std::array<std::vector<float>, 3> flow_wind, flow_loop;
    bool flow_round = false;
    float flow_fastest = 0.0f;
    size_t flow_sweeps = 0;

    Stencil st;
    std::array<float, 18> weights{};  // the stencil's lowered weights, as the layout hands them out
    Branches branches;

    // layout
    std::vector<float> scale, layer_scale;
    std::vector<uint8_t> type;
    std::vector<uint16_t> cut;
    SpecialList specials;
    std::vector<int64_t> layer_cells;
    std::array<std::vector<float>, 6> keep, push;
    std::vector<uint64_t> probe_start{0}, probe_cells;
    std::vector<double> probe_weights;
    std::vector<uint8_t> probe_found, probe_in_sight;
    size_t air = 0, boundary_cells = 0;
    bool open = false;
    std::array<bool, 6> open_sides{};
    // Fitted walls: the mesh as given, each special's volume (cells^3), whether the source
    // lies on a face.
    bool fitted_walls = false;
    fitted::Mesh mesh;
    std::vector<double> special_volume;
    bool source_on_face = false;
    // The fit's rows and cell kinds, kept for cut_cells() (the binding's, native_module.cpp:
    // 1614-1632) and its counts.
    std::vector<int64_t> cut_rows;
    std::vector<double> cut_air_volume, cut_open_area, cut_wall_area, cut_sliver;
    std::vector<uint8_t> cut_kind;
    fitted::FitStats cut_stats;

    // triangle bins for crossing queries
    size_t bin = 8, bx = 0, by = 0, bz = 0;
    std::vector<uint32_t> bin_start, bin_faces;
    std::vector<std::array<uint32_t, 6>> face_bins;  // each face's bins: from x, y, z, to x, y, z
    // Which faces a query has met (each thread's own: the bins list a face more than once).
    struct Stamps {
        std::vector<uint32_t> mark;
        uint32_t now = 0;
    };
    mutable Stamps serial;

    V3 centre_at(long long x, long long y, long long z) const {
        return {origin.x + (static_cast<double>(x) + 0.5) * cell, origin.y + (static_cast<double>(y) + 0.5) * cell,
                origin.z + (static_cast<double>(z) + 0.5) * cell};
    }
    V3 centre_of(size_t i) const {
        return centre_at(static_cast<long long>(i % nx), static_cast<long long>((i / nx) % ny),
                         static_cast<long long>(i / sz));
    }
    size_t neighbour(size_t i, int k) const {
        return static_cast<size_t>(static_cast<ptrdiff_t>(i) + st.d[static_cast<size_t>(k)].offset);
    }
    bool edge(size_t i) const {
        const size_t x = i % nx, y = (i / nx) % ny, z = i / sz;
        return x == 0 || y == 0 || z == 0 || x + 1 == nx || y + 1 == ny || z + 1 == nz;
    }
    double cell_scale(size_t i) const { return scale.empty() ? 1.0 : static_cast<double>(scale[i]); }
    // Cells from i to the grid's edge, the nearest way (0: on the edge).
    size_t depth(size_t i) const {
        const size_t x = i % nx, y = (i / nx) % ny, z = i / sz;
        return std::min({x, y, z, nx - 1 - x, ny - 1 - y, nz - 1 - z});
    }
    // Whether a face cuts the link from i along k (k < st.count): the bit of the link's end
    // whose direction owns it.
    bool is_cut(size_t i, int k) const {
        const Direction &d = st.d[static_cast<size_t>(k)];
        if (d.plus_bit >= 0) return (cut[i] >> d.plus_bit) & 1u;
        return (cut[neighbour(i, k)] >> st.d[static_cast<size_t>(d.opposite)].plus_bit) & 1u;
    }
    // Whether the neighbour of the cell at (x, y, z) along k lies on the grid's edge (a step
    // past the grid counts as on it).
    bool neighbour_on_edge(size_t x, size_t y, size_t z, int k) const {
        const int *s = st.d[static_cast<size_t>(k)].step;
        const long long to[3] = {static_cast<long long>(x) + s[0], static_cast<long long>(y) + s[1],
                                 static_cast<long long>(z) + s[2]};
        const long long size[3] = {static_cast<long long>(nx), static_cast<long long>(ny), static_cast<long long>(nz)};
        for (int a = 0; a < 3; ++a)
            if (to[a] <= 0 || to[a] >= size[a] - 1) return true;
        return false;
    }
    bool blocked_at(size_t i, int k, size_t x, size_t y, size_t z) const {
        return neighbour_on_edge(x, y, z, k) || is_cut(i, k);
    }
    LinkMask links_at(size_t i, size_t x, size_t y, size_t z) const {
        LinkMask m;
        for (int k = 0; k < st.count; ++k) {
            if (neighbour_on_edge(x, y, z, k)) {
                m.edge_far |= 1u << k;
                m.blocked |= 1u << k;
            } else if (is_cut(i, k)) {
                m.blocked |= 1u << k;
            }
        }
        return m;
    }
    LinkMask links_of(size_t i) const { return links_at(i, i % nx, (i / nx) % ny, i / sz); }

    // Runs work(t, z_lo, z_hi) for slabs of planes t = 0 .. T - 1 on T threads (T returned
    // by slabs()); the first slab's exception, by order of slabs, is thrown again.
    size_t slabs() const { return std::max<size_t>(1, std::min<size_t>(static_cast<size_t>(threads), nz / 2)); }
    template <typename Work>
    void in_slabs(size_t T, Work work) const {
        std::vector<std::exception_ptr> errors(T);
        auto run = [&](size_t t) {
            try {
                work(t, nz * t / T, nz * (t + 1) / T);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < T; ++t) pool.emplace_back(run, t);
        run(0);
        for (auto &th : pool) th.join();
        for (const auto &e : errors)
            if (e) std::rethrow_exception(e);
    }


    // --- geometry -------------------------------------------------------------------

    // The segment from a to b against triangle f: 0 no meeting, 1 a firm crossing at
    // fraction t, 2 a meeting within the margins (t set). Throws where an end lies on the
    // face's plane within the guard, next to the face.
    int meet(const Triangle &f, const V3 &a, const V3 &b, double &t) const {
        const double sa = dot(f.normal, a - f.v0), sb = dot(f.normal, b - f.v0);
        const auto near_box = [&](const V3 &p) {
            return p.x >= f.lo.x - guard && p.x <= f.hi.x + guard && p.y >= f.lo.y - guard && p.y <= f.hi.y + guard &&
                   p.z >= f.lo.z - guard && p.z <= f.hi.z + guard;
        };
        if ((std::abs(sa) <= guard && near_box(a)) || (std::abs(sb) <= guard && near_box(b))) {
            throw Degenerate("degenerate geometry: a cell centre, the source or a probe lies within " +
                             std::to_string(guard) + " m of a face's plane next to the face");
        }
        if ((sa > 0.0) == (sb > 0.0)) return 0;
        t = sa / (sa - sb);
        const V3 p = a + (b - a) * t;
        // Barycentric coordinates of p in the face's plane.
        const V3 h = cross(f.normal, f.e2);
        const double det = dot(f.e1, h);
        if (!(std::abs(det) > 0.0)) return 0;
        const V3 s = p - f.v0;
        const double u = dot(s, h) / det;
        const double v = dot(s, cross(f.e1, f.normal)) / det;
        const double w = 1.0 - u - v;
        const double least = std::min({u, v, w});
        if (least > kBarycentric) return 1;
        if (least < -kBarycentric) return 0;
        return 2;
    }

    template <typename Visit>
    void each_face_near_segment(const V3 &a, const V3 &b, Stamps &st, Visit visit) const {
        const double lo[3] = {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
        const double hi[3] = {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
        const size_t dims[3] = {bx, by, bz};
        const double o[3] = {origin.x, origin.y, origin.z};
        size_t from[3], to[3];
        for (int k = 0; k < 3; ++k) {
            const double f = std::floor((lo[k] - o[k]) / cell / static_cast<double>(bin)) - 1.0;
            const double l = std::floor((hi[k] - o[k]) / cell / static_cast<double>(bin)) + 1.0;
            from[k] = static_cast<size_t>(std::clamp(f, 0.0, static_cast<double>(dims[k] - 1)));
            to[k] = static_cast<size_t>(std::clamp(l, 0.0, static_cast<double>(dims[k] - 1)));
        }
        if (st.mark.size() != tri.size()) {
            st.mark.assign(tri.size(), 0u);
            st.now = 0;
        }
        if (++st.now == 0) {
            std::fill(st.mark.begin(), st.mark.end(), 0u);
            st.now = 1;
        }
        for (size_t z = from[2]; z <= to[2]; ++z)
            for (size_t y = from[1]; y <= to[1]; ++y)
                for (size_t x = from[0]; x <= to[0]; ++x) {
                    const size_t c = x + bx * (y + by * z);
                    for (uint32_t q = bin_start[c]; q < bin_start[c + 1]; ++q) {
                        const uint32_t f = bin_faces[q];
                        if (st.mark[f] == st.now) continue;
                        st.mark[f] = st.now;
                        visit(f);
                    }
                }
    }

    // The face nearest a that the segment from a to b crosses, or -1 (the reference's
    // crossing, fdtd.hpp:3082-3257, where no tie touches it).
    int64_t crossing(const V3 &a, const V3 &b, Stamps &st) const {
        Hit firm;
        struct Loose {
            double t;
            int64_t face;
        };
        Loose loose[16];
        int nloose = 0;
        bool twice = false;
        each_face_near_segment(a, b, st, [&](uint32_t f) {
            double t = 0;
            const int r = meet(tri[f], a, b, t);
            if (r == 1) {
                if (firm.face < 0 || t < firm.t) {
                    if (firm.face >= 0 && std::abs(t - firm.t) <= 1e-12 && tri[f].material != tri[firm.face].material)
                        twice = true;
                    firm = {t, static_cast<int64_t>(f)};
                } else if (std::abs(t - firm.t) <= 1e-12 && tri[f].material != tri[firm.face].material) {
                    twice = true;
                }
            } else if (r == 2) {
                if (nloose == 16) throw Degenerate("degenerate geometry: a segment passes through many faces' edges");
                loose[nloose++] = {t, static_cast<int64_t>(f)};
            }
        });
        if (twice) throw Degenerate("degenerate geometry: a segment crosses two faces of different materials at one point");
        for (int q = 0; q < nloose; ++q) {
            const Loose &l = loose[q];
            if (firm.face < 0) throw Degenerate("degenerate geometry: a segment passes through a face's edge");
            const bool same_place = std::abs(l.t - firm.t) <= 1e-9 && tri[l.face].material == tri[firm.face].material;
            const bool farther = l.t > firm.t + 1e-9;
            if (!same_place && !farther) {
                throw Degenerate("degenerate geometry: a segment passes through a face's edge next to its nearest crossing");
            }
        }
        return firm.face;
    }

    int64_t crossing(const V3 &a, const V3 &b) const { return crossing(a, b, serial); }
    bool sees(const V3 &p, const V3 &q) const { return crossing(p, q) < 0; }

    // Whether p lies on a face: refused (the reference's leans are not taken).
    void require_off_faces(const V3 &p, const char *what) const {
        each_face_near_segment(p, p, serial, [&](uint32_t f) {
            const Triangle &t = tri[f];
            const double s = dot(t.normal, p - t.v0);
            if (std::abs(s) > guard) return;
            const V3 h = cross(t.normal, t.e2);
            const double det = dot(t.e1, h);
            if (!(std::abs(det) > 0.0)) return;
            const V3 d = p - t.v0;
            const double u = dot(d, h) / det, v = dot(d, cross(t.e1, t.normal)) / det;
            if (std::min({u, v, 1.0 - u - v}) >= -kBarycentric) {
                throw std::invalid_argument(std::string(what) +
                                            " lies on a surface of the model: move it off the surface into the air");
            }
        });
    }

    // --- the grid ------------------------------------------------------------------

    void make_bins() {
        bx = (nx + bin - 1) / bin;
        by = (ny + bin - 1) / bin;
        bz = (nz + bin - 1) / bin;
        const size_t nb = bx * by * bz;
        std::vector<uint32_t> count(nb + 1, 0);
        const auto range = [&](const Triangle &t, size_t *from, size_t *to) {
            const double lo[3] = {t.lo.x, t.lo.y, t.lo.z}, hi[3] = {t.hi.x, t.hi.y, t.hi.z};
            const double o[3] = {origin.x, origin.y, origin.z};
            const size_t dims[3] = {bx, by, bz};
            for (int k = 0; k < 3; ++k) {
                const double f = std::floor((lo[k] - o[k]) / cell / static_cast<double>(bin)) - 1.0;
                const double l = std::floor((hi[k] - o[k]) / cell / static_cast<double>(bin)) + 1.0;
                from[k] = static_cast<size_t>(std::clamp(f, 0.0, static_cast<double>(dims[k] - 1)));
                to[k] = static_cast<size_t>(std::clamp(l, 0.0, static_cast<double>(dims[k] - 1)));
            }
        };
        face_bins.assign(tri.size(), {});
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<uint32_t> fill;
            if (pass == 1) {
                bin_start.assign(nb + 1, 0);
                for (size_t c = 0; c < nb; ++c) bin_start[c + 1] = bin_start[c] + count[c];
                bin_faces.assign(bin_start[nb], 0);
                fill.assign(bin_start.begin(), bin_start.end() - 1);
            }
            for (size_t f = 0; f < tri.size(); ++f) {
                size_t from[3], to[3];
                range(tri[f], from, to);
                for (int a = 0; a < 3; ++a) {
                    face_bins[f][static_cast<size_t>(a)] = static_cast<uint32_t>(from[a]);
                    face_bins[f][static_cast<size_t>(a) + 3] = static_cast<uint32_t>(to[a]);
                }
                for (size_t z = from[2]; z <= to[2]; ++z)
                    for (size_t y = from[1]; y <= to[1]; ++y)
                        for (size_t x = from[0]; x <= to[0]; ++x) {
                            const size_t c = x + bx * (y + by * z);
                            if (pass == 0) {
                                ++count[c];
                            } else {
                                bin_faces[fill[c]++] = static_cast<uint32_t>(f);
                            }
                        }
            }
        }
        serial.mark.assign(tri.size(), 0u);
    }

    // Cut links (fdtd.hpp:446-477): every face against the + half's links of the cells
    // whose centres lie within a cell of its box; threads own slabs of planes.
    void cut_links() {
        cut.assign(n, 0);
        const int T = std::max(1, threads);
        std::vector<std::vector<std::pair<size_t, int>>> unsure(static_cast<size_t>(T));
        std::vector<std::string> errors(static_cast<size_t>(T));
        auto work = [&](int t) {
            const size_t z_lo = nz * static_cast<size_t>(t) / static_cast<size_t>(T);
            const size_t z_hi = nz * static_cast<size_t>(t + 1) / static_cast<size_t>(T);
            try {
                for (const Triangle &f : tri) {
                    size_t from[3], to[3];
                    bool empty = false;
                    const double lo[3] = {f.lo.x, f.lo.y, f.lo.z}, hi[3] = {f.hi.x, f.hi.y, f.hi.z};
                    const double o[3] = {origin.x, origin.y, origin.z};
                    const size_t size[3] = {nx, ny, nz};
                    for (int a = 0; a < 3; ++a) {
                        const double first = std::ceil((lo[a] - cell - o[a]) / cell - 0.5) - 1.0;
                        const double last = std::floor((hi[a] + cell - o[a]) / cell - 0.5) + 1.0;
                        if (last < 0.0 || first > static_cast<double>(size[a] - 1)) {
                            empty = true;
                            break;
                        }
                        from[a] = static_cast<size_t>(std::max(first, 0.0));
                        to[a] = std::min(static_cast<size_t>(last), size[a] - 1);
                    }
                    if (empty) continue;
                    from[2] = std::max(from[2], z_lo);
                    if (z_hi == 0 || to[2] < z_lo) continue;
                    to[2] = std::min(to[2], z_hi - 1);
                    for (size_t z = from[2]; z <= to[2] && z < z_hi; ++z)
                        for (size_t y = from[1]; y <= to[1]; ++y)
                            for (size_t x = from[0]; x <= to[0]; ++x) {
                                const size_t i = x + y * sy + z * sz;
                                const V3 c = centre_at(static_cast<long long>(x), static_cast<long long>(y),
                                                       static_cast<long long>(z));
                                for (int k = 0; k < st.count; ++k) {
                                    if (st.d[static_cast<size_t>(k)].plus_bit < 0) continue;
                                    const long long tx = static_cast<long long>(x) + st.d[static_cast<size_t>(k)].step[0];
                                    const long long ty = static_cast<long long>(y) + st.d[static_cast<size_t>(k)].step[1];
                                    const long long tz = static_cast<long long>(z) + st.d[static_cast<size_t>(k)].step[2];
                                    if (tx < 0 || ty < 0 || tz < 0 || tx >= static_cast<long long>(nx) ||
                                        ty >= static_cast<long long>(ny) || tz >= static_cast<long long>(nz))
                                        continue;
                                    double frac = 0;
                                    const int r = meet(f, c, centre_at(tx, ty, tz), frac);
                                    if (r == 1) {
                                        cut[i] = static_cast<uint16_t>(cut[i] | (1u << st.d[static_cast<size_t>(k)].plus_bit));
                                    } else if (r == 2) {
                                        unsure[static_cast<size_t>(t)].push_back({i, k});
                                    }
                                }
                            }
                }
            } catch (const std::exception &e) {
                errors[static_cast<size_t>(t)] = e.what();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
        for (const auto &e : errors)
            if (!e.empty()) throw Degenerate(e);
        // A link that meets a face only within the margins is cut or not by the tie rules:
        // refused, unless another face cuts it firmly.
        for (const auto &list : unsure)
            for (const auto &[i, k] : list)
                if (!((cut[i] >> st.d[static_cast<size_t>(k)].plus_bit) & 1u)) throw Degenerate("degenerate geometry: a link passes through a face's edge");
    }

    // --- readings ----------------------------------------------------------------------

    // The cells round p (up to eight) whose centres p sees, weighed trilinearly and scaled
    // to sum to one; only cells `admit` takes (seen_cells, fdtd.hpp:3305-3365, for p on no
    // face). Empty when p lies within a cell of the grid's edge or what is seen weighs
    // under 1e-9.
    template <typename Admit>
    std::vector<std::pair<size_t, double>> trilinear(const V3 &p, Admit admit) const {
        std::vector<std::pair<size_t, double>> out;
        const double at[3] = {(p.x - origin.x) / cell - 0.5, (p.y - origin.y) / cell - 0.5,
                              (p.z - origin.z) / cell - 0.5};
        const size_t size[3] = {nx, ny, nz};
        size_t low[3];
        double side[3][2];  // per axis: the weights of the lower and upper cell
        for (int a = 0; a < 3; ++a) {
            if (!(at[a] >= 1.0 && at[a] <= static_cast<double>(size[a]) - 3.0)) return out;
            const double f = std::floor(at[a]);
            low[a] = static_cast<size_t>(f);
            side[a][1] = at[a] - f;
            side[a][0] = 1.0 - side[a][1];
        }
        double total = 0.0;
        for (unsigned corner = 0; corner < 8; ++corner) {
            const unsigned ux = corner & 1u, uy = (corner >> 1) & 1u, uz = (corner >> 2) & 1u;
            const double w = side[0][ux] * side[1][uy] * side[2][uz];
            if (w <= 0.0) continue;
            const size_t i = (low[0] + ux) + (low[1] + uy) * sy + (low[2] + uz) * sz;
            if (!admit(i) || !sees(p, centre_of(i))) continue;
            out.emplace_back(i, w);
            total += w;
        }
        if (total < 1e-9) {
            out.clear();
            return out;
        }
        for (auto &cw : out) cw.second /= total;
        return out;
    }

    std::vector<std::pair<size_t, double>> visible(const V3 &p, bool air_only) const {
        return trilinear(p, [&](size_t i) { return !air_only || type[i] != kSolid; });
    }

    // The reading of p (reading_cells, fdtd.hpp:2361-2488, for p on no face). Trilinear
    // where `cubic` is false. Otherwise the 4 x 4 x 4 block round p: its cells that are air
    // and in sight of p; the planes of each axis kept while each holds such a cell among
    // the other axes' kept planes (the largest such set); per axis, cubic weights where all
    // four planes are kept, linear where the middle two are, 1 where one of them is; every
    // cell of the box so chosen must be in sight and linked to its next ones in the box.
    // Trilinear wherever any of that fails.
    std::vector<std::pair<size_t, double>> reading(const V3 &p, bool cubic) const {
        if (!cubic) return visible(p, true);
        const double at[3] = {(p.x - origin.x) / cell - 0.5, (p.y - origin.y) / cell - 0.5,
                              (p.z - origin.z) / cell - 0.5};
        const size_t size[3] = {nx, ny, nz};
        long long first[3];
        double below_frac[3], w[3][kReach];
        for (int a = 0; a < 3; ++a) {
            const double f = std::floor(at[a]);
            const auto below = static_cast<long long>(f);
            if (!(at[a] >= 1.0 && at[a] < static_cast<double>(size[a]) - 3.0 && below >= 1)) return visible(p, true);
            first[a] = below - 1;
            below_frac[a] = at[a] - f;
            const double x = at[a] - static_cast<double>(first[a]);
            for (int j = 0; j < kReach; ++j) w[a][j] = lagrange4(x, j);
        }
        const auto index = [&](int x, int y, int z) {
            return static_cast<size_t>(first[0] + x) + static_cast<size_t>(first[1] + y) * sy +
                   static_cast<size_t>(first[2] + z) * sz;
        };
        uint64_t seen = 0;
        for (int z = 0; z < kReach; ++z)
            for (int y = 0; y < kReach; ++y)
                for (int x = 0; x < kReach; ++x) {
                    const size_t i = index(x, y, z);
                    if ((type[i] == kInterior || type[i] == kSpecial) && sees(p, centre_of(i))) seen |= block_bit(x, y, z);
                }
        unsigned kept[3] = {0xFu, 0xFu, 0xFu};
        for (bool shrunk = true; shrunk;) {
            shrunk = false;
            for (int a = 0; a < 3; ++a) {
                const uint64_t across = seen & block_slab((a + 1) % 3, kept[(a + 1) % 3]) & block_slab((a + 2) % 3, kept[(a + 2) % 3]);
                for (int j = 0; j < kReach; ++j) {
                    if (((kept[a] >> j) & 1u) && !(across & block_slab(a, 1u << j))) {
                        kept[a] &= ~(1u << j);
                        shrunk = true;
                    }
                }
            }
        }
        int lo[3], hi[3];  // the box, per axis [lo, hi)
        for (int a = 0; a < 3; ++a) {
            if (kept[a] == 0xFu) {
                lo[a] = 0;
                hi[a] = kReach;
            } else if ((kept[a] & 0x6u) == 0x6u) {
                lo[a] = 1;
                hi[a] = 3;
                w[a][1] = 1.0 - below_frac[a];
                w[a][2] = below_frac[a];
            } else if (kept[a] & 0x6u) {
                lo[a] = (kept[a] & 0x2u) ? 1 : 2;
                hi[a] = lo[a] + 1;
                w[a][lo[a]] = 1.0;
            } else {
                return visible(p, true);
            }
        }
        std::vector<std::pair<size_t, double>> out;
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x) {
                    const size_t i = index(x, y, z);
                    const bool linked = !(x + 1 < hi[0] && is_cut(i, 1)) && !(y + 1 < hi[1] && is_cut(i, 3)) &&
                                        !(z + 1 < hi[2] && is_cut(i, 5));
                    if (!(seen & block_bit(x, y, z)) || !linked) return visible(p, true);
                    const double wi = w[0][x] * w[1][y] * w[2][z];
                    if (wi != 0.0) out.emplace_back(i, wi);
                }
        return out;
    }

    // Every cell the queue's cells (already air) reach by links not blocked; the set does
    // not depend on the order. By spans along x: a span is marked whole (each end extended
    // through open links to cells not yet air), then every link of its cells off the x axis
    // is followed, each cell it reaches opening a span of its own. The grid is read row by
    // row, with one stack entry a span.
    void flood(std::vector<size_t> &queue) {
        struct Span {
            size_t row, x0, x1;  // row = i / nx; cells x0 .. x1 of it
        };
        std::vector<Span> spans;
        // Marks the span through air cell i (already marked) and pushes it.
        const auto open_span = [&](size_t i) {
            const size_t row = i / nx, x = i - row * nx, y = row % ny, z = row / ny;
            size_t x0 = x, x1 = x;
            while (!blocked_at(row * nx + x0, 0, x0, y, z) && type[row * nx + x0 - 1] == kSolid) {
                type[row * nx + x0 - 1] = kInterior;
                --x0;
            }
            while (!blocked_at(row * nx + x1, 1, x1, y, z) && type[row * nx + x1 + 1] == kSolid) {
                type[row * nx + x1 + 1] = kInterior;
                ++x1;
            }
            spans.push_back({row, x0, x1});
        };
        for (const size_t i : queue) open_span(i);
        queue.clear();
        while (!spans.empty()) {
            const Span sp = spans.back();
            spans.pop_back();
            const size_t y = sp.row % ny, z = sp.row / ny;
            for (int k = 2; k < st.count; ++k) {
                for (size_t x = sp.x0; x <= sp.x1; ++x) {
                    const size_t i = sp.row * nx + x;
                    if (blocked_at(i, k, x, y, z)) continue;
                    const size_t j = neighbour(i, k);
                    if (type[j] != kSolid) continue;
                    type[j] = kInterior;
                    open_span(j);
                }
            }
        }
    }

    // --- wall cells ----------------------------------------------------------------------

    // A contact: a wall a blocked link's boundary weight goes to, across the cell's link
    // along `dir`.
    struct Contact {
        int dir;
        double weight;
    };

    // A wall cell's contacts in direction order (fdtd.hpp:602-614): a blocked axial link
    // meets the wall across itself; a blocked diagonal meets the walls across those of its
    // two axial sides that are blocked (the lower axis first), or, where neither is, the
    // wall across itself; each contact carries the blocked link's boundary weight.
    int contacts(const LinkMask &m, Contact *out) const {
        int count = 0;
        for (int k = 0; k < st.count; ++k) {
            if (!((m.blocked >> k) & 1u)) continue;
            const Direction &d = st.d[static_cast<size_t>(k)];
            if (k < 6) {
                out[count++] = {k, d.wall};
                continue;
            }
            const bool a = (m.blocked >> d.side_a) & 1u, b = (m.blocked >> d.side_b) & 1u;
            if (a) out[count++] = {d.side_a, d.wall};
            if (b) out[count++] = {d.side_b, d.wall};
            if (!a && !b) out[count++] = {k, d.wall};
        }
        return count;
    }

    // What cell i's contact along `dir` meets (fdtd.hpp:581-593): the grid's edge where the
    // neighbour lies on it; else the material of the face nearest i that the link between
    // the two centres crosses, looked for from i, then from the neighbour; else nothing.
    int32_t wall_of(size_t i, int dir, Stamps &stamps) const {
        const size_t j = neighbour(i, dir);
        if (edge(j)) return kEdgeWall;
        int64_t face = crossing(centre_of(i), centre_of(j), stamps);
        if (face < 0) face = crossing(centre_of(j), centre_of(i), stamps);
        if (face < 0) return kNoWall;
        return static_cast<int32_t>(face_material[static_cast<size_t>(face)]);
    }

    // Boundary cells (fdtd.hpp:546-625), in stages over slabs of planes on the threads:
    //  1. an interior cell with a blocked link is a wall cell (mark 1);
    //  2. an interior cell that is not, next to a wall cell across a link whose folded
    //     weight differs from the stencil's, is marked too (mark 2): averaging changes it;
    //  3. each marked cell, in grid order: its link weights, each the mean of its folded
    //     weight and its neighbour's towards it (the stencil's where the neighbour is
    //     unmarked air or in the matched layer; 0 towards solid and edge cells); its wall
    //     contacts, resolved; the wall term g, the sum of boundary weight times admittance
    //     over the contacts in order, plus, per material met (in the order first met), its
    //     contacts' weight times each of its branches' gain; one branch state per branch of
    //     each material met. The slabs' lists are joined in slab order: specials ascend.
    void wall_cells(size_t T) {
        std::vector<uint8_t> mark(n, 0);
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (size_t z = z0; z < z1; ++z)
                for (size_t y = 0; y < ny; ++y)
                    for (size_t x = 0; x < nx; ++x) {
                        const size_t i = x + y * sy + z * sz;
                        if (type[i] == kInterior && links_at(i, x, y, z).blocked) mark[i] = 1;
                    }
        });
        std::vector<uint8_t> reached(n, 0);
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            std::array<double, 18> w{};
            for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                if (mark[i] != 1) continue;
                fold(st, links_of(i), w.data());
                for (int k = 0; k < st.count; ++k) {
                    const size_t j = neighbour(i, k);
                    if (w[static_cast<size_t>(k)] != st.d[static_cast<size_t>(k)].exact && type[j] == kInterior &&
                        !mark[j]) {
                        __atomic_store_n(&reached[j], static_cast<uint8_t>(1), __ATOMIC_RELAXED);
                    }
                }
            }
        });
        for (size_t i = 0; i < n; ++i)
            if (reached[i]) mark[i] = 2;

        struct Part {
            SpecialList list;
            size_t air = 0;
            bool open = false;
            std::array<bool, 6> sides{};
        };
        std::vector<Part> parts(T);
        in_slabs(T, [&](size_t t, size_t z0, size_t z1) {
            Part &part = parts[t];
            Stamps stamps;
            std::array<double, 18> mine{}, theirs{}, weight{};
            std::array<Contact, 36> touch{};
            std::array<int32_t, 36> met{};        // materials met, in the order first met
            std::array<double, 36> met_weight{};  // and their contacts' weights
            std::vector<SpecialList::State> states;
            for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                if (type[i] != kInterior) continue;
                ++part.air;
                if (!mark[i]) continue;
                const LinkMask m = links_of(i);
                fold(st, m, mine.data());
                weight.fill(0.0);
                for (int k = 0; k < st.count; ++k) {
                    const auto ku = static_cast<size_t>(k);
                    const size_t j = neighbour(i, k);
                    if (((m.edge_far >> k) & 1u) || type[j] == kSolid) continue;
                    double toward = st.d[ku].exact;
                    if (type[j] == kInterior && mark[j]) {
                        fold(st, links_of(j), theirs.data());
                        toward = theirs[static_cast<size_t>(st.d[ku].opposite)];
                    }
                    weight[ku] = 0.5 * (mine[ku] + toward);
                }
                const int count = contacts(m, touch.data());
                double g = 0.0;
                size_t kinds = 0;
                for (int c = 0; c < count; ++c) {
                    const Contact &ct = touch[static_cast<size_t>(c)];
                    const int32_t wall = wall_of(i, ct.dir, stamps);
                    if (wall == kNoWall) continue;
                    if (wall == kEdgeWall) {
                        g += ct.weight * edge_admittance;
                        part.open = true;
                        const size_t j = neighbour(i, ct.dir);
                        const size_t at[3] = {j % nx, (j / nx) % ny, j / sz}, size[3] = {nx, ny, nz};
                        for (size_t a = 0; a < 3; ++a) {
                            if (at[a] == 0) part.sides[2 * a] = true;
                            if (at[a] + 1 >= size[a]) part.sides[2 * a + 1] = true;
                        }
                        continue;
                    }
                    g += ct.weight * beta[static_cast<size_t>(wall)];
                    size_t kind = 0;
                    while (kind < kinds && met[kind] != wall) ++kind;
                    if (kind == kinds) {
                        met[kinds] = wall;
                        met_weight[kinds++] = 0.0;
                    }
                    met_weight[kind] += ct.weight;
                }
                states.clear();
                const double r = std::sqrt(cell_scale(i));
                for (size_t kind = 0; kind < kinds; ++kind) {
                    for (const int32_t jb : branches.of_material[static_cast<size_t>(met[kind])]) {
                        g += met_weight[kind] * (0.5 / branches.a[static_cast<size_t>(jb)]);
                        states.push_back({jb, static_cast<float>(r * met_weight[kind])});
                    }
                }
                part.list.append(i, weight.data(), st.count, cell_scale(i), g, states.data(), states.size());
            }
        });
        for (const Part &part : parts) {
            for (const int64_t i : part.list.cell) type[static_cast<size_t>(i)] = kSpecial;
            specials.extend(part.list);
            air += part.air;
            open = open || part.open;
            for (size_t k = 0; k < 6; ++k) open_sides[k] = open_sides[k] || part.sides[k];
        }
        specials.sorted = specials.size();
        boundary_cells = specials.size();
    }

    // The matched layer: every air cell within `layer` cells of the grid's edge, which no
    // cut link may reach (fdtd.hpp:521-542), slab by slab on the threads (each cell's class
    // its own; the cells listed in grid order), and its coefficients.
    void matched_layer(size_t T) {
        std::vector<std::vector<int64_t>> found(T);
        std::vector<std::array<bool, 6>> sides(T);
        in_slabs(T, [&](size_t t, size_t z0, size_t z1) {
            std::array<bool, 6> reach{};
            const size_t size[3] = {nx, ny, nz};
            for (size_t z = z0; z < z1; ++z)
                for (size_t y = 0; y < ny; ++y)
                    for (size_t x = 0; x < nx; ++x) {
                        const size_t i = x + y * sy + z * sz;
                        if (type[i] == kSolid) continue;
                        if (std::min({x, y, z, nx - 1 - x, ny - 1 - y, nz - 1 - z}) > layer) continue;
                        for (int k = 0; k < 6; ++k) {
                            if (!neighbour_on_edge(x, y, z, k) && is_cut(i, k)) {
                                throw std::invalid_argument("the model reaches into the matched layer: a face cuts a "
                                                            "link between two of its cells");
                            }
                        }
                        type[i] = kLayer;
                        found[t].push_back(static_cast<int64_t>(i));
                        const size_t at[3] = {x, y, z};
                        for (size_t a = 0; a < 3; ++a) {
                            if (at[a] <= layer) reach[2 * a] = true;
                            if (at[a] + 1 + layer >= size[a]) reach[2 * a + 1] = true;
                        }
                    }
            sides[t] = reach;
        });
        for (size_t t = 0; t < T; ++t) {
            layer_cells.insert(layer_cells.end(), found[t].begin(), found[t].end());
            air += found[t].size();
            if (!found[t].empty()) open = true;
            for (size_t k = 0; k < 6; ++k) open_sides[k] = open_sides[k] || sides[t][k];
        }
        layer_profile();
    }

    // The matched layer's coefficients along each axis, and its cells' speeds.
    void layer_profile() {
        if (layer_cells.empty()) return;
        const double courant = std::sqrt(static_cast<double>(st.d[0].weight));
        const size_t size[3] = {nx, ny, nz};
        for (size_t a = 0; a < 3; ++a) {
            LayerAxis ax = layer_axis(size[a], layer, layer_damping, courant);
            keep[2 * a] = std::move(ax.keep_centre);
            keep[2 * a + 1] = std::move(ax.keep_face);
            push[2 * a] = std::move(ax.push_centre);
            push[2 * a + 1] = std::move(ax.push_face);
        }
        if (!scale.empty())
            for (const int64_t c : layer_cells) layer_scale.push_back(scale[static_cast<size_t>(c)]);
    }

    // The source's cells (fdtd.hpp:627-638): none of the cells round it in the matched
    // layer; each cell of its reading made special (with the stencil's weights where it was
    // interior) and fed s w, s its speed factor.
    void place_source(const V3 &src) {
        for (const auto &cw : visible(src, true))
            if (type[cw.first] == kLayer) throw std::invalid_argument("the source lies in the matched layer");
        std::array<double, 18> exact{};
        for (int k = 0; k < 18; ++k) exact[static_cast<size_t>(k)] = st.d[static_cast<size_t>(k)].exact;
        for (const auto &[i, w] : reading(src, true)) {
            size_t k = specials.index_of(i);
            if (k == npos) {
                type[i] = kSpecial;
                specials.append(i, exact.data(), st.count, cell_scale(i), 0.0, nullptr, 0);
                k = specials.size() - 1;
            }
            specials.source[k] += static_cast<float>(cell_scale(i) * w);
        }
    }

    // The probes' readings (fdtd.hpp:639-690): trilinear within three cells of the source,
    // cubic elsewhere; whether each probe's listener (the first of its group) sees it.
    void place_probes(const std::vector<V3> &probes, const V3 &src, size_t per_listener) {
        for (const V3 &p : probes) {
            const bool near = norm(p - src) < (kNear - 1e-9) * cell;
            const auto cells = reading(p, !near);
            for (const auto &[i, w] : cells) {
                probe_cells.push_back(i);
                probe_weights.push_back(w);
            }
            probe_start.push_back(probe_cells.size());
            probe_found.push_back(cells.empty() ? 0 : 1);
        }
        for (size_t k = 0; k < probes.size(); ++k) {
            const size_t listener = k - k % per_listener;
            probe_in_sight.push_back(k == listener || sees(probes[listener], probes[k]) ? 1 : 0);
        }
    }

    void check_arguments(size_t probe_count, size_t per_listener) const {
        if (points != 7 && points != 19) {
            throw std::invalid_argument("a stencil of 7 or 19 points is needed, not " + std::to_string(points));
        }
        if (per_listener == 0 || probe_count % per_listener != 0) {
            throw std::invalid_argument("the probes must come in whole groups of " + std::to_string(per_listener) +
                                        " (one group a listener)");
        }
        if (layer > 0 && points != 7) throw std::invalid_argument("a matched layer works with the 7-point stencil only");
        if (nx < 3 || ny < 3 || nz < 3) throw std::invalid_argument("the grid needs at least 3 cells along each axis");
        if (!(cell > 0.0)) throw std::invalid_argument("the cell size must be positive");
        if (2 * layer + 3 > std::min({nx, ny, nz})) {
            throw std::invalid_argument("the matched layer is too thick: it leaves no cells of the grid inside it");
        }
        if (!(layer_damping >= 0.0)) throw std::invalid_argument("the matched layer's damping must not be negative");
        for (size_t f = 0; f < face_material.size(); ++f) {
            if (face_material[f] < 0 || static_cast<size_t>(face_material[f]) >= materials) {
                throw std::invalid_argument("face " + std::to_string(f) + " has no material");
            }
        }
        for (size_t m = 0; m < materials; ++m) {
            if (!(beta[m] >= 0.0)) {
                throw std::invalid_argument("material " + std::to_string(m) + " has a negative admittance");
            }
        }
    }

    // --- moving air ---------------------------------------------------------------------

    bool moving(size_t i) const { return type[i] == kInterior || type[i] == kSpecial; }

    // Moving air (set_up_flow, fdtd.hpp:1680-1737): each moving cell's shifts a step along x
    // and y (under half a cell), and which of its four neighbours in the plane are air across
    // open links. Plain where along each axis the air moves a cell its two neighbours are air
    // moving as it does (or not moving at all): then each cell's mask says, per axis (4 bits
    // each, x low), which of its row's cells one and two away move with it. Otherwise the
    // wind round the solids (set_up_round).
    void set_up_flow(const float *fx, const float *fy) {
        flow_x.assign(n, 0.0f);
        flow_y.assign(n, 0.0f);
        flow_mask.assign(n, 0);
        std::vector<uint8_t> near(n, 0);  // bit k: the neighbour along k (-x, +x, -y, +y)
        for (size_t i = 0; i < n; ++i) {
            if (!moving(i)) continue;
            if (!(std::isfinite(fx[i]) && std::isfinite(fy[i]) && std::abs(fx[i]) < 0.5f && std::abs(fy[i]) < 0.5f))
                throw std::invalid_argument("the air must move less than half a cell per step");
            flow_x[i] = fx[i];
            flow_y[i] = fy[i];
            flow_fastest = std::max({flow_fastest, std::abs(fx[i]), std::abs(fy[i])});
            const size_t x = i % nx, y = (i / nx) % ny, z = i / sz;
            for (int k = 0; k < 4; ++k)
                if (!blocked_at(i, k, x, y, z) && type[neighbour(i, k)] != kSolid) near[i] |= static_cast<uint8_t>(1u << k);
        }
        const auto moves = [&](size_t from, int k) { return ((near[from] >> k) & 1u) && moving(neighbour(from, k)); };
        bool plain = true;
        for (size_t i = 0; i < n && plain; ++i) {
            if (!moving(i)) continue;
            for (int axis = 0; axis < 2 && plain; ++axis) {
                const std::vector<float> &f = axis == 0 ? flow_x : flow_y;
                if (f[i] == 0.0f) continue;
                for (int k = 2 * axis; k < 2 * axis + 2; ++k) {
                    const size_t j = neighbour(i, k);
                    plain = plain && ((near[i] >> k) & 1u) && (!moving(j) || f[j] == f[i]);
                }
            }
        }
        if (!plain) {
            set_up_round();
            return;
        }
        for (size_t i = 0; i < n; ++i) {
            if (!moving(i)) continue;
            unsigned m = 0;
            for (int axis = 0; axis < 2; ++axis) {
                const int lo = 2 * axis, hi = lo + 1;
                const unsigned at = 4u * static_cast<unsigned>(axis);
                const bool below = moves(i, lo), above = moves(i, hi);
                if (below) m |= 1u << at;
                if (above) m |= 2u << at;
                if (below && moves(neighbour(i, lo), lo)) m |= 4u << at;
                if (above && moves(neighbour(i, hi), hi)) m |= 8u << at;
            }
            flow_mask[i] = static_cast<uint8_t>(m);
        }
    }

    // The wind round the solids (set_up_round, fdtd.hpp:1739-1935): the given wind made free
    // of sources by a potential lambda solved on the faces open to the wind,
    //   sum over a cell's open faces of k (lambda - lambda') = -(the given wind's net flux),
    // k a face's weight (the two cells' mean given speed, or the moving one's beside still
    // air), by conjugate gradients scaled by the diagonal, in double, to a net flux of 1e-9 of
    // the fastest given speed. Every sum runs plane by plane along z, each plane in index
    // order, then over the planes in order. Then the faces' winds, the loops (a twelfth of
    // each three moving cells in a row, halved), and which cells the air moves.
    void set_up_round() {
        const size_t stride[3] = {1, sy, sz};
        const size_t least = layer > 0 ? 1 : 2;
        std::vector<float> speed(n, 0.0f);
        std::vector<uint8_t> faces(n, 0);  // bit d: the face towards +d is open to the wind
        const std::vector<float> *given[2] = {&flow_x, &flow_y};
        const size_t T = slabs();
        // Faces before any weight: a face's flag is read from the cell below it, which may
        // lie in another slab.
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                if (type[i] == kSolid || depth(i) < least) continue;
                speed[i] = moving(i) ? std::hypot(flow_x[i], flow_y[i]) : 0.0f;
                const size_t x = i % nx, y = (i / nx) % ny, z = i / sz;
                for (int d = 0; d < 3; ++d) {
                    const size_t j = i + stride[d];
                    if (type[j] == kSolid || depth(j) < least || blocked_at(i, 2 * d + 1, x, y, z)) continue;
                    if (moving(i) || moving(j)) faces[i] |= static_cast<uint8_t>(1u << d);
                }
            }
        });
        const auto mean = [&](const std::vector<float> &f, size_t i, size_t j) {
            const double sum = static_cast<double>(f[i]) + static_cast<double>(f[j]);
            return moving(i) && moving(j) ? 0.5 * sum : sum;
        };
        const auto through = [&](size_t i, int d) { return d < 2 ? mean(*given[d], i, i + stride[d]) : 0.0; };
        std::vector<double> lambda(n, 0.0), residual(n, 0.0), direction(n, 0.0), image(n, 0.0);
        std::vector<float> diagonal(n, 0.0f);
        // The solve runs on T threads, each on its slab of planes, in step: term(i) summed (or
        // its largest) over the moving cells of each plane in index order, the planes' sums
        // in two buffers taken in turn, then every thread takes the planes in order to the
        // same total. One wait a sum: a thread can be one sum ahead, writing the other buffer.
        std::vector<double> planes[2] = {std::vector<double>(nz, 0.0), std::vector<double>(nz, 0.0)};
        std::atomic<size_t> arrived{0}, generation{0};
        const auto wait_all = [&]() {
            const size_t g = generation.load(std::memory_order_acquire);
            if (arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == T) {
                arrived.store(0, std::memory_order_relaxed);
                generation.fetch_add(1, std::memory_order_release);
                return;
            }
            while (generation.load(std::memory_order_acquire) == g) std::this_thread::yield();
        };
        const size_t most = 1000 + 50 * (nx + ny + nz);
        size_t sweeps = 0;
        double top = 0.0;  // the fastest given speed
        in_slabs(T, [&](size_t t, size_t z0, size_t z1) {
            int turn_of_buffer = 0;
            const auto over = [&](bool largest, auto &&term) {
                std::vector<double> &plane_sums = planes[turn_of_buffer];
                turn_of_buffer ^= 1;
                for (size_t z = z0; z < z1; ++z) {
                    double plane = 0.0;
                    for (size_t i = z * sz; i < (z + 1) * sz; ++i) {
                        if (!moving(i)) continue;
                        const double v = term(i);
                        plane = largest ? std::max(plane, v) : plane + v;
                    }
                    plane_sums[z] = plane;
                }
                wait_all();
                double total = 0.0;
                for (size_t z = 0; z < nz; ++z) total = largest ? std::max(total, plane_sums[z]) : total + plane_sums[z];
                return total;
            };
            const double fastest_given = over(true, [&](size_t i) {
                double flux = 0.0, weights = 0.0;
                for (int d = 0; d < 3; ++d) {
                    const size_t s = stride[d];
                    if ((faces[i] >> d) & 1u) {
                        flux += through(i, d);
                        weights += mean(speed, i, i + s);
                    }
                    if ((faces[i - s] >> d) & 1u) {
                        flux -= through(i - s, d);
                        weights += mean(speed, i - s, i);
                    }
                }
                residual[i] = weights > 0.0 ? -flux : 0.0;
                diagonal[i] = static_cast<float>(weights);
                return static_cast<double>(speed[i]);
            });
            const auto scaled = [&](size_t i) {
                return diagonal[i] > 0.0f ? residual[i] / static_cast<double>(diagonal[i]) : 0.0;
            };
            double carried = over(false, [&](size_t i) {
                direction[i] = scaled(i);
                return residual[i] * direction[i];
            });
            const double enough = 1e-9 * fastest_given;
            double left = over(true, [&](size_t i) { return std::abs(residual[i]); });
            size_t mine = 0;
            while (left > enough) {
                if (++mine > most) throw std::runtime_error("the wind round the model's solids did not settle");
                // (direction's neighbours in the slabs beside are final: the last sum's
                // wait came after every thread wrote its own.)
                const double bent = over(false, [&](size_t i) {
                    double sum = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const size_t s = stride[d];
                        if ((faces[i] >> d) & 1u) sum += mean(speed, i, i + s) * (direction[i] - direction[i + s]);
                        if ((faces[i - s] >> d) & 1u) sum += mean(speed, i - s, i) * (direction[i] - direction[i - s]);
                    }
                    image[i] = sum;
                    return direction[i] * sum;
                });
                if (!(bent > 0.0)) break;
                const double step = carried / bent;
                left = over(true, [&](size_t i) {
                    lambda[i] += step * direction[i];
                    residual[i] -= step * image[i];
                    return std::abs(residual[i]);
                });
                const double next = over(false, [&](size_t i) { return residual[i] * scaled(i); });
                const double turn = next / carried;
                carried = next;
                over(false, [&](size_t i) {
                    direction[i] = scaled(i) + turn * direction[i];
                    return 0.0;
                });
            }
            if (t == 0) {
                sweeps = mine;
                top = fastest_given;
            }
        });
        flow_sweeps = sweeps;
        // The faces' winds and the fastest of them.
        for (int d = 0; d < 3; ++d) flow_wind[static_cast<size_t>(d)].assign(n, 0.0f);
        std::vector<double> &fastest_of = planes[0];
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (size_t z = z0; z < z1; ++z) {
                double plane = 0.0;
                for (size_t i = z * sz; i < (z + 1) * sz; ++i) {
                    for (int d = 0; d < 3; ++d) {
                        if (!((faces[i] >> d) & 1u)) continue;
                        const size_t j = i + stride[d];
                        const double u = through(i, d) - mean(speed, i, j) * (lambda[j] - lambda[i]);
                        flow_wind[static_cast<size_t>(d)][i] = static_cast<float>(u);
                        plane = std::max(plane, std::abs(u));
                    }
                }
                fastest_of[z] = plane;
            }
        });
        double fastest = 0.0;
        for (size_t z = 0; z < nz; ++z) fastest = std::max(fastest, fastest_of[z]);
        flow_fastest = static_cast<float>(fastest);
        // A wind nowhere a millionth of the given one (a closed room's) is none.
        if (static_cast<double>(flow_fastest) <= 1e-6 * top) {
            flow_fastest = 0.0f;
            for (auto &w : flow_wind) std::fill(w.begin(), w.end(), 0.0f);
        }
        if (!(flow_fastest < 0.5f)) {
            throw std::invalid_argument("the air must move less than half a cell per step, and the wind round the "
                                        "model's solids moves " + std::to_string(flow_fastest) + " at its fastest");
        }
        flow_x.clear();
        flow_y.clear();
        flow_x.shrink_to_fit();
        flow_y.shrink_to_fit();
        for (int d = 0; d < 3; ++d) flow_loop[static_cast<size_t>(d)].assign(n, 0.0f);
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (int d = 0; d < 3; ++d) {
                std::vector<float> &loop = flow_loop[static_cast<size_t>(d)];
                const std::vector<float> &w = flow_wind[static_cast<size_t>(d)];
                const size_t s = stride[d];
                for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                    if (!((faces[i] >> d) & 1u) || !((faces[i + s] >> d) & 1u)) continue;
                    if (!(moving(i) && moving(i + s) && moving(i + 2 * s))) continue;
                    loop[i] = (w[i] + w[i + s]) * (1.0f / 24.0f);
                }
            }
        });
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                if (!moving(i) || depth(i) < 2) continue;
                bool moved = false, border = false;
                for (int d = 0; d < 3; ++d) {
                    const size_t s = stride[d];
                    const std::vector<float> &w = flow_wind[static_cast<size_t>(d)], &l = flow_loop[static_cast<size_t>(d)];
                    moved = moved || w[i] != 0.0f || w[i - s] != 0.0f || l[i] != 0.0f || l[i - s] != 0.0f ||
                            l[i - 2 * s] != 0.0f;
                    border = border || (w[i] != 0.0f && !moving(i + s)) || (w[i - s] != 0.0f && !moving(i - s));
                }
                flow_mask[i] = moved ? (border ? 2 : 1) : 0;
            }
        });
        flow_round = true;
    }

    static constexpr const char *kNoSight = "no cell centre of the wave grid is in sight of the source";

    // The passes over the whole grid on the current GPU (setup_device.hip): the cut links,
    // the air, the matched layer and the wall cells, with the host passes' results and
    // refusals, in the host's order (a guard fault or a link through a face's edge, then
    // no air, then the layer, then the wall cells' refusal at the lowest cell).
    void device_passes(const V3 &src) {
        asdev::Input in;
        asdev::Grid &g = in.g;
        g.origin[0] = origin.x;
        g.origin[1] = origin.y;
        g.origin[2] = origin.z;
        g.cell = cell;
        g.guard = guard;
        g.barycentric = kBarycentric;
        g.nx = static_cast<uint32_t>(nx);
        g.ny = static_cast<uint32_t>(ny);
        g.nz = static_cast<uint32_t>(nz);
        g.bin = static_cast<uint32_t>(bin);
        g.bx = static_cast<uint32_t>(bx);
        g.by = static_cast<uint32_t>(by);
        g.bz = static_cast<uint32_t>(bz);
        g.directions = static_cast<uint32_t>(st.count);
        g.layer = static_cast<uint32_t>(layer);
        g.edge_admittance = edge_admittance;
        for (size_t k = 0; k < 18; ++k) {
            const Direction &d = st.d[k];
            for (int a = 0; a < 3; ++a) g.step[k][a] = static_cast<int8_t>(d.step[a]);
            g.plus_bit[k] = static_cast<int8_t>(d.plus_bit);
            g.opposite[k] = static_cast<int8_t>(d.opposite);
            g.side_a[k] = static_cast<int8_t>(d.side_a);
            g.side_b[k] = static_cast<int8_t>(d.side_b);
            g.exact[k] = d.exact;
            g.wall[k] = d.wall;
        }
        if (nx >= (1u << 31) || ny >= (1u << 31) || nz >= (1u << 31) || layer >= (1u << 31))
            throw std::invalid_argument("the grid is too large for the device set-up");
        in.tri.resize(tri.size());
        const auto put = [](const V3 &v, double *o) {
            o[0] = v.x;
            o[1] = v.y;
            o[2] = v.z;
        };
        for (size_t f = 0; f < tri.size(); ++f) {
            asdev::Tri &t = in.tri[f];
            put(tri[f].v0, t.v0);
            put(tri[f].e1, t.e1);
            put(tri[f].e2, t.e2);
            put(tri[f].normal, t.n);
            put(tri[f].lo, t.lo);
            put(tri[f].hi, t.hi);
            t.material = face_material[f];
            for (int a = 0; a < 3; ++a) {
                t.bin_from[a] = face_bins[f][static_cast<size_t>(a)];
                t.bin_to[a] = face_bins[f][static_cast<size_t>(a) + 3];
            }
        }
        in.bin_start = bin_start;
        in.bin_faces = bin_faces;
        in.seeds.assign(n, 0);
        size_t seeds = 0;
        for (const auto &cw : visible(src, false)) {
            if (in.seeds[cw.first]) continue;
            in.seeds[cw.first] = 1;
            ++seeds;
        }
        in.scale = scale.empty() ? nullptr : scale.data();
        in.beta = beta;
        in.branch_start.assign(1, 0);
        for (const auto &list : branches.of_material) {
            in.branch_of.insert(in.branch_of.end(), list.begin(), list.end());
            in.branch_start.push_back(static_cast<uint32_t>(in.branch_of.size()));
        }
        in.branch_a = branches.a;

        asdev::Output out;
        asdev::set_up(in, out);
        if (std::getenv("AS_FDTD_TIMING") != nullptr) {
            std::fprintf(stderr,
                         "as_setup_create: device cut links %.1f ms, flood %.1f ms in %zu rounds, layer and marks "
                         "%.1f ms, wall cells %.1f ms\n",
                         1e3 * out.seconds[0], 1e3 * out.seconds[1], out.rounds, 1e3 * out.seconds[2],
                         1e3 * out.seconds[3]);
        }
        flood_rounds = out.rounds;
        const std::string guard_fault = "degenerate geometry: a cell centre, the source or a probe lies within " +
                                        std::to_string(guard) + " m of a face's plane next to the face";
        switch (out.fault) {
            case asdev::kGuard: throw Degenerate(guard_fault);
            case asdev::kEdgeLink: throw Degenerate("degenerate geometry: a link passes through a face's edge");
            case asdev::kEdgeOnly: throw Degenerate("degenerate geometry: a segment passes through a face's edge");
            default: break;
        }
        if (seeds == 0) throw std::invalid_argument(kNoSight);
        switch (out.fault) {
            case asdev::kLayerReach:
                throw std::invalid_argument("the model reaches into the matched layer: a face cuts a link between two "
                                            "of its cells");
            case asdev::kManyEdges: throw Degenerate("degenerate geometry: a segment passes through many faces' edges");
            case asdev::kTwice:
                throw Degenerate("degenerate geometry: a segment crosses two faces of different materials at one point");
            case asdev::kEdgeNear:
                throw Degenerate("degenerate geometry: a segment passes through a face's edge next to its nearest "
                                 "crossing");
            default: break;
        }
        cut = std::move(out.cut);
        type = std::move(out.type);
        layer_cells = std::move(out.layer_cells);
        air += layer_cells.size() + out.interior;
        open = open || out.open;
        for (size_t k = 0; k < 6; ++k) open_sides[k] = open_sides[k] || ((out.sides >> k) & 1u);
        layer_profile();
        SpecialList &sp = specials;
        sp.cell = std::move(out.cell);
        sp.link_start = std::move(out.link_start);
        sp.state_start = std::move(out.state_start);
        sp.link_direction = std::move(out.link_direction);
        sp.link_weight = std::move(out.link_weight);
        sp.coef_a = std::move(out.coef_a);
        sp.coef_b = std::move(out.coef_b);
        sp.coef_inv = std::move(out.coef_inv);
        sp.source.reserve(sp.cell.size() + 64);
        sp.source.assign(sp.cell.size(), 0.0f);
        sp.state_branch = std::move(out.state_branch);
        sp.state_weight = std::move(out.state_weight);
        sp.sorted = sp.size();
        boundary_cells = sp.size();
    }

    // --- fitted walls ----------------------------------------------------------------------

    // The sliver rule's factor for a sliver f cells thick, lambda = 1 + f:
    // 3 lambda / ((lambda + 1/2)(lambda + 1)), at most 1 (fdtd.hpp:950-955).
    static double sliver_factor(double f) {
        const double lambda = 1.0 + f;
        return 3.0 * lambda / ((lambda + 0.5) * (lambda + 1.0));
    }

    // A cell's reading on the fit (the fitted stencil, fdtd.hpp:1455-1510): the eight cells
    // round the located point that `keep` takes and whose centres it reaches crossing no
    // face, weighed trilinearly; along an axis where the point is tied to a centre, the
    // cells past it only where none before it is seen; scaled to sum to one. `edge` is set
    // where the point lies within a cell of the grid's edge.
    template <typename Keep>
    std::vector<std::pair<size_t, double>> seen_fitted(const fitted::Fit &fit, const fitted::Located &L, Keep keep,
                                                       bool *at_edge = nullptr, std::vector<int64_t> *crossed = nullptr) const {
        std::vector<std::pair<size_t, double>> outv;
        const long long size[3] = {static_cast<long long>(nx), static_cast<long long>(ny), static_cast<long long>(nz)};
        for (int a = 0; a < 3; ++a) {
            const long long j = L.j[a];
            if (!(j >= 1 && (j < size[a] - 3 || (j == size[a] - 3 && L.tied[a])))) {
                if (at_edge) *at_edge = true;
                return outv;
            }
        }
        struct Seen {
            size_t cell;
            double w;
            int past;
        };
        Seen seen[8];
        int count = 0, least = 0;
        for (int corner = 0; corner < 8; ++corner) {
            double w = 1.0;
            int past = 0;
            size_t i = 0, stride = 1;
            for (int a = 0; a < 3; ++a) {
                const int up = (corner >> a) & 1;
                if (!L.tied[a]) w *= up ? L.lo[a] : L.hi[a];
                else if (up) past += 1 << (2 * (a + 1));
                i += static_cast<size_t>(L.j[a] + up) * stride;
                stride *= a == 0 ? nx : (a == 1 ? ny : 1);
            }
            if (w <= 0.0 || !keep(i)) continue;
            int64_t c[3];
            fit.centre_q(i, c);
            const int64_t f = fit.way(L.q, c);
            if (f >= 0) {
                if (crossed) crossed->push_back(f);
                continue;
            }
            if (count == 0 || past < least) least = past;
            seen[count++] = {i, w, past};
        }
        double total = 0.0;
        for (int k = 0; k < count; ++k) {
            if (seen[k].past != least) continue;
            outv.emplace_back(seen[k].cell, seen[k].w);
            total += seen[k].w;
        }
        for (auto &cw : outv) cw.second /= total;
        return outv;
    }

    // The reading on the fit (fdtd.hpp:1515-1598): the cubic block of the located point's
    // four planes per axis where its cells are air, seen and linked in the block, else
    // linear or single along an axis, else trilinear (seen_fitted). Each weight factor
    // x - m is formed as lo - (m - 1), so a tied point weighs 0 exactly.
    std::vector<std::pair<size_t, double>> reading_fitted(const fitted::Fit &fit, const fitted::Located &L, bool cubic) const {
        const auto visible = [&]() { return seen_fitted(fit, L, [&](size_t i) { return fit.kind[i] != fitted::kSolid; }); };
        if (!cubic) return visible();
        const long long size[3] = {static_cast<long long>(nx), static_cast<long long>(ny), static_cast<long long>(nz)};
        long long first[3];
        double w[3][kReach];
        for (int a = 0; a < 3; ++a) {
            const long long j = L.j[a];
            if (!(j >= 1 && j < size[a] - 3)) return visible();
            first[a] = j - 1;
            for (int k = 0; k < kReach; ++k) {
                double v = 1.0;
                for (int m = 0; m < kReach; ++m)
                    if (m != k) v *= (L.lo[a] - static_cast<double>(m - 1)) / static_cast<double>(k - m);
                w[a][k] = v;
            }
        }
        const auto index = [&](int x, int y, int z) {
            return static_cast<size_t>(first[0] + x) + static_cast<size_t>(first[1] + y) * sy +
                   static_cast<size_t>(first[2] + z) * sz;
        };
        const auto clear = [&](size_t i, size_t j) {
            int64_t a[3], b[3];
            fit.centre_q(i, a);
            fit.centre_q(j, b);
            return fit.way(a, b) < 0;
        };
        uint64_t own = 0;
        for (int z = 0; z < kReach; ++z)
            for (int y = 0; y < kReach; ++y)
                for (int x = 0; x < kReach; ++x) {
                    const size_t i = index(x, y, z);
                    if (fit.kind[i] != fitted::kAir) continue;
                    int64_t c[3];
                    fit.centre_q(i, c);
                    if (fit.way(L.q, c) < 0) own |= block_bit(x, y, z);
                }
        unsigned kept[3] = {0xFu, 0xFu, 0xFu};
        for (bool shrunk = true; shrunk;) {
            shrunk = false;
            for (int a = 0; a < 3; ++a) {
                const uint64_t across = own & block_slab((a + 1) % 3, kept[(a + 1) % 3]) & block_slab((a + 2) % 3, kept[(a + 2) % 3]);
                for (int j = 0; j < kReach; ++j) {
                    if (((kept[a] >> j) & 1u) && !(across & block_slab(a, 1u << j))) {
                        kept[a] &= ~(1u << j);
                        shrunk = true;
                    }
                }
            }
        }
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            if (kept[a] == 0xFu) {
                lo[a] = 0;
                hi[a] = kReach;
            } else if ((kept[a] & 0x6u) == 0x6u) {
                lo[a] = 1;
                hi[a] = 3;
                w[a][1] = L.hi[a];
                w[a][2] = L.lo[a];
            } else if (kept[a] & 0x6u) {
                lo[a] = (kept[a] & 0x2u) ? 1 : 2;
                hi[a] = lo[a] + 1;
                w[a][lo[a]] = 1.0;
            } else {
                return visible();
            }
        }
        std::vector<std::pair<size_t, double>> outv;
        for (int z = lo[2]; z < hi[2]; ++z)
            for (int y = lo[1]; y < hi[1]; ++y)
                for (int x = lo[0]; x < hi[0]; ++x) {
                    const size_t i = index(x, y, z);
                    if (!(own & block_bit(x, y, z)) || (x + 1 < hi[0] && !clear(i, i + 1)) ||
                        (y + 1 < hi[1] && !clear(i, i + sy)) || (z + 1 < hi[2] && !clear(i, i + sz))) {
                        return visible();
                    }
                    const double wi = w[0][x] * w[1][y] * w[2][z];
                    if (wi != 0.0) outv.emplace_back(i, wi);
                }
        return outv;
    }

    // The layout on fitted walls (set_up_fitted, fdtd.hpp:970-1217; add_special_fitted
    // 1221-1264): cells' kinds from the fit; every row, and every plain cell the sliver rule
    // reaches, a special cell with coefficients of its own (a row whose coefficients are a
    // plain cell's to the bit stays plain); the source's and probes' cells from the fit.
    void build_fitted(const V3 &src, const std::vector<V3> &probes, size_t per_listener) {
        SetupClock clock;
        fitted::GridSpec g;
        g.origin[0] = origin.x;
        g.origin[1] = origin.y;
        g.origin[2] = origin.z;
        g.cell = cell;
        g.n[0] = nx;
        g.n[1] = ny;
        g.n[2] = nz;
        g.layer = layer;
        const double source[3] = {src.x, src.y, src.z};
        const fitted::Fit fit(mesh, face_material.data(), materials, g, source, threads);
        clock.mark("fit");
        type.assign(n, kSolid);
        air = 0;
        for (size_t i = 0; i < n; ++i) {
            if (fit.kind[i] == fitted::kAir) {
                type[i] = kInterior;
                ++air;
            } else if (fit.kind[i] == fitted::kLayer) {
                type[i] = kLayer;
                layer_cells.push_back(static_cast<int64_t>(i));
                ++air;
            }
        }
        open_sides = fit.open_sides;
        open = false;
        for (const bool b : open_sides) open = open || b;
        layer_profile();

        const std::vector<int64_t> &rows = fit.rows;
        const size_t columns = materials + 1;
        const auto row_of = [&rows](size_t i) -> size_t {
            const auto it = std::lower_bound(rows.begin(), rows.end(), static_cast<int64_t>(i));
            return it != rows.end() && *it == static_cast<int64_t>(i) ? static_cast<size_t>(it - rows.begin()) : npos;
        };
        // The sliver rule: a link through the face opposite a sliver of thickness f > 0 that
        // a row owns is scaled by w(f) on both sides; where both its cells' rules touch it,
        // by the product of their factors.
        struct Touch {
            size_t key;  // 3 (the link's lower cell) + its axis
            double factor;
        };
        std::vector<Touch> touches;
        for (size_t r = 0; r < rows.size(); ++r) {
            const auto o = static_cast<size_t>(rows[r]);
            for (int k = 0; k < 6; ++k) {
                const double f = fit.sliver[6 * r + static_cast<size_t>(k)];
                if (!(f > 0.0)) continue;
                const size_t m = neighbour(o, st.d[static_cast<size_t>(k)].opposite);
                if (edge(m) || type[m] == kSolid || type[m] == kLayer) continue;
                touches.push_back({3 * std::min(o, m) + static_cast<size_t>(k / 2), sliver_factor(f)});
            }
        }
        std::sort(touches.begin(), touches.end(), [](const Touch &x, const Touch &y) {
            return x.key != y.key ? x.key < y.key : x.factor < y.factor;
        });
        std::vector<std::pair<size_t, double>> factors;
        std::vector<size_t> far_cells;
        for (size_t t = 0; t < touches.size();) {
            const size_t key = touches[t].key;
            double product = 1.0;
            for (; t < touches.size() && touches[t].key == key; ++t) product *= touches[t].factor;
            factors.emplace_back(key, product);
            const size_t axis = key % 3, lower = key / 3, upper = lower + (axis == 0 ? 1 : (axis == 1 ? sy : sz));
            for (const size_t c : {lower, upper})
                if (row_of(c) == npos) far_cells.push_back(c);
        }
        std::sort(far_cells.begin(), far_cells.end());
        far_cells.erase(std::unique(far_cells.begin(), far_cells.end()), far_cells.end());
        const auto factor_of = [&factors](size_t i, size_t j, int axis) {
            const size_t key = 3 * std::min(i, j) + static_cast<size_t>(axis);
            const auto it = std::lower_bound(factors.begin(), factors.end(), std::make_pair(key, -1.0));
            return it != factors.end() && it->first == key ? it->second : 1.0;
        };
        const auto room = [this](size_t i) { return type[i] == kInterior || type[i] == kSpecial; };
        // A cell's open area towards direction k, after the sliver rule.
        const auto area_towards = [&](size_t i, int k) {
            const size_t r = row_of(i);
            const size_t j = neighbour(i, k);
            double a = 0.0;
            if (r != npos) a = fit.open_area[18 * r + static_cast<size_t>(k)];
            else if (k < 6 && type[i] != kSolid && !edge(j) && type[j] != kSolid) a = fit.face_area(i, k / 2);
            if (k < 6 && a > 0.0 && room(i) && room(j)) a *= factor_of(i, j, k / 2);
            return a;
        };
        const double exact = st.d[0].exact, plane = st.d[0].wall, half_g = exact / 2.0;
        struct FittedCell {
            double volume = 1.0, air_volume = 1.0, wall = 0.0;
            std::array<double, 18> area{};
            std::vector<std::pair<int32_t, double>> states;
        };
        const auto make = [&](size_t i) {
            FittedCell c;
            const size_t r = row_of(i);
            c.air_volume = r != npos ? fit.air_volume[r] : fit.cell_volume(i);
            for (int k = 0; k < 18; ++k) c.area[static_cast<size_t>(k)] = area_towards(i, k);
            if (r != npos) {
                for (size_t m = 0; m + 1 < columns; ++m) {
                    const double W = fit.wall_area[columns * r + m];
                    if (!(W > 0.0)) continue;
                    const double weight = plane * W;
                    c.wall += weight * beta[m];
                    for (const int32_t j : branches.of_material[m]) {
                        c.wall += weight * (0.5 / branches.a[static_cast<size_t>(j)]);
                        c.states.emplace_back(j, weight);
                    }
                }
                const double edge_wall = fit.wall_area[columns * r + columns - 1];
                if (edge_wall > 0.0) c.wall += plane * edge_wall * edge_admittance;
            }
            double total = 0.0;
            for (const double a : c.area) total += a;
            c.volume = std::max(c.air_volume, half_g * total);
            return c;
        };
        // Makes cell i special with c's coefficients, unless `spare` and they are a plain
        // cell's to the bit; returns whether it did.
        const auto add = [&](size_t i, const FittedCell &c, bool spare) {
            const double s = cell_scale(i), r = std::sqrt(s) / c.volume;
            float weight[18] = {};
            double total = 0.0;
            for (int k = 0; k < 18; ++k) {
                const double a = c.area[static_cast<size_t>(k)];
                if (!(a > 0.0)) continue;
                weight[k] = lowered(s * (exact * a / c.volume));
                total += static_cast<double>(weight[k]);
            }
            const double centre = 2.0 - total;
            const float coef_a = down(centre);
            const auto coef_b = static_cast<float>(1.0 - r * c.wall);
            const auto coef_inv = static_cast<float>(1.0 / (1.0 + r * c.wall));
            if (spare && scale.empty() && c.states.empty() && coef_a == st.centre && coef_b == 1.0f && coef_inv == 1.0f) {
                bool plain = true;
                for (int k = 0; k < 18 && plain; ++k) plain = weight[k] == (k < 6 ? weights[static_cast<size_t>(k)] : 0.0f);
                if (plain) return false;
            }
            type[i] = kSpecial;
            std::vector<SpecialList::State> states;
            for (const auto &[j, w] : c.states) states.push_back({j, static_cast<float>(r * w)});
            specials.append_given(i, c.area.data(), weight, coef_a, coef_b, coef_inv, states.data(), states.size());
            special_volume.push_back(c.volume);
            return true;
        };
        std::vector<size_t> candidates(far_cells);
        for (const int64_t i : rows) candidates.push_back(static_cast<size_t>(i));
        std::sort(candidates.begin(), candidates.end());
        for (const size_t i : candidates) add(i, make(i), true);
        specials.sorted = specials.size();
        boundary_cells = specials.size();
        clock.mark("fitted specials");

        // The source: each of its cells fed w / V.
        const fitted::Located S = fit.locate(&source[0]);
        source_on_face = fit.on_face(S.q);
        for (const auto &cw : seen_fitted(fit, S, [&](size_t i) { return fit.kind[i] != fitted::kSolid; }))
            if (fit.kind[cw.first] == fitted::kLayer) throw std::invalid_argument("fitted walls: the source is in the matched layer");
        const auto feed = reading_fitted(fit, S, true);
        if (feed.empty()) {
            bool at_edge = false;
            std::vector<int64_t> crossed;
            seen_fitted(fit, S, [&](size_t i) { return fit.kind[i] != fitted::kSolid; }, &at_edge, &crossed);
            if (at_edge) {
                throw std::invalid_argument("fitted walls: the source lies within " + std::to_string(kReach / 2) +
                                            " cells of the grid's edge, where the solver has no cells round it: move it "
                                            "inward or enlarge the grid.");
            }
            throw std::invalid_argument("fitted walls: no cell centre round the source can be reached from it without "
                                        "crossing the model: the source lies in a space narrower than a cell. Move it, "
                                        "or use a finer grid.");
        }
        std::array<double, 18> stencil_weights{};
        for (int k = 0; k < 18; ++k) stencil_weights[static_cast<size_t>(k)] = st.d[static_cast<size_t>(k)].exact;
        for (const auto &[i, w] : feed) {
            size_t k = specials.index_of(i);
            if (k == npos) {
                if (row_of(i) != npos) {
                    add(i, make(i), false);
                } else {
                    type[i] = kSpecial;
                    specials.append(i, stencil_weights.data(), st.count, cell_scale(i), 0.0, nullptr, 0);
                    special_volume.push_back(fit.cell_volume(i));
                }
                k = specials.size() - 1;
            }
            specials.source[k] += static_cast<float>(cell_scale(i) * w / special_volume[k]);
        }
        // The probes: cubic, or trilinear within three cells of the source; whether each is
        // in sight of its listener.
        std::vector<fitted::Located> at;
        for (const V3 &p : probes) {
            const bool close = norm(p - src) < (kNear - 1e-9) * cell;
            const double q[3] = {p.x, p.y, p.z};
            at.push_back(fit.locate(q));
            const auto cells = reading_fitted(fit, at.back(), !close);
            for (const auto &[i, w] : cells) {
                probe_cells.push_back(i);
                probe_weights.push_back(w);
            }
            probe_start.push_back(probe_cells.size());
            probe_found.push_back(cells.empty() ? 0 : 1);
        }
        for (size_t k = 0; k < probes.size(); ++k) {
            const size_t l = k - k % per_listener;
            if (k == l) {
                probe_in_sight.push_back(1);
            } else if (!fit.in_grid_box(at[l].q) || !fit.in_grid_box(at[k].q)) {
                probe_in_sight.push_back(0);
            } else {
                probe_in_sight.push_back(fit.way(at[l].q, at[k].q) < 0 ? 1 : 0);
            }
        }
        clock.mark("fitted source, probes");
        cut_rows = fit.rows;
        cut_air_volume = fit.air_volume;
        cut_open_area = fit.open_area;
        cut_wall_area = fit.wall_area;
        cut_sliver = fit.sliver;
        cut_kind = fit.kind;
        cut_stats = fit.stats;

        // The links' self-check: each link stands for one area from both of its cells,
        // reaches only air, is a whole face towards the matched layer, and a cell's links
        // weigh at most 2 s together.
        for (size_t k = 0; k < specials.size(); ++k) {
            const auto i = static_cast<size_t>(specials.cell[k]);
            double sum = 0.0;
            for (uint64_t l = specials.link_start[k]; l < specials.link_start[k + 1]; ++l) {
                const int d = specials.link_direction[l];
                const size_t j = neighbour(i, d);
                sum += static_cast<double>(specials.link_weight[l]);
                const double here = area_towards(i, d), there = area_towards(j, st.d[static_cast<size_t>(d)].opposite);
                const char *fault = nullptr;
                if (edge(j) || type[j] == kSolid) fault = "a link reaches a cell without air";
                else if (std::abs(here - there) > 1e-12 * std::max(here, there)) fault = "a link stands for different areas from its two cells";
                else if (type[j] == kLayer && (d >= 6 || std::abs(here - fit.face_area(i, d / 2)) > 1e-12))
                    fault = "a link to the matched layer is not a whole face";
                if (fault) throw std::logic_error(std::string("fitted walls: self-check of the links failed at cell ") + std::to_string(i) + ": " + fault);
            }
            if (sum > 2.0 * cell_scale(i)) {
                throw std::logic_error("fitted walls: self-check of the links failed at cell " + std::to_string(i) +
                                       ": its links weigh more than 2");
            }
        }
    }

    void build(const double *mrq, const int64_t *branch_material, size_t branch_count, const V3 &src,
               const std::vector<V3> &probes, size_t per_listener, const float *speed, const float *fx,
               const float *fy) {
        check_arguments(probes.size(), per_listener);
        if ((fx == nullptr) != (fy == nullptr)) throw std::invalid_argument("the air's motion needs both of its components");
        branches = make_branches(mrq, branch_material, branch_count, materials);
        st = make_stencil(points, static_cast<ptrdiff_t>(sy), static_cast<ptrdiff_t>(sz));
        for (size_t k = 0; k < 18; ++k) weights[k] = st.d[k].weight;
        if (speed != nullptr) {
            scale.resize(n);
            for (size_t i = 0; i < n; ++i) {
                const double r = static_cast<double>(speed[i]);
                if (!(r > 0.0 && r <= 1.0)) {
                    throw std::invalid_argument("cell " + std::to_string(i) +
                                                ": its speed of sound over the grid's must lie in (0, 1]");
                }
                scale[i] = static_cast<float>(r * r);
            }
        }
        if (fitted_walls) {
            // The reference's own refusals (fdtd.hpp:429-445), then this set-up's.
            if (points != 7) {
                throw std::invalid_argument("fitted walls: they need the 7-point stencil, and the 19-point one was "
                                            "asked for: set the stencil to 7, or take walls on cell faces");
            }
            if (fx != nullptr) {
                throw std::invalid_argument("fitted walls: they are not available in moving air: leave the air's "
                                            "motion out, or take walls on cell faces");
            }
            if (speed != nullptr) {
                throw std::invalid_argument("fitted walls: they take one speed of sound for the whole grid, and a "
                                            "speed per cell was given: leave the speed of each cell out, or take "
                                            "walls on cell faces");
            }
            if (on_device) throw std::invalid_argument("fitted walls: the fit runs on the host only");
            build_fitted(src, probes, per_listener);
            return;
        }
        SetupClock clock;
        make_bins();
        clock.mark("stencil, bins");
        require_off_faces(src, "the source");
        for (const V3 &p : probes) require_off_faces(p, "a probe");
        if (on_device) {
            device_passes(src);
            clock.mark("device passes");
            place_source(src);
            place_probes(probes, src, per_listener);
            clock.mark("source, probes");
            if (fx != nullptr) set_up_flow(fx, fy);
            clock.mark("moving air");
            return;
        } else {
            cut_links();
            clock.mark("cut links");

            // Air: the cells the source sees, and all they reach (fdtd.hpp:498-517).
            type.assign(n, kSolid);
            std::vector<size_t> queue;
            for (const auto &cw : visible(src, false)) {
                if (type[cw.first] != kSolid) continue;
                type[cw.first] = kInterior;
                queue.push_back(cw.first);
            }
            if (queue.empty()) throw std::invalid_argument(kNoSight);
            flood(queue);
            clock.mark("air");
        }
        const size_t T = slabs();
        if (layer > 0) matched_layer(T);
        clock.mark("layer");
        wall_cells(T);
        clock.mark("marks, specials");
        place_source(src);
        place_probes(probes, src, per_listener);
        clock.mark("source, probes");
        if (fx != nullptr) set_up_flow(fx, fy);
        clock.mark("moving air");
    }
};

thread_local std::string g_setup_error;

// Threads for `threads` 0: OMP_NUM_THREADS where it is set to a positive count (the share
// of the machine the process is given), else every hardware thread.
int default_threads() {
    if (const char *env = std::getenv("OMP_NUM_THREADS")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v > 0 && v <= 4096) return static_cast<int>(v);
    }
    return static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
}

}  // namespace

struct as_setup {
    Setup s;
};

extern "C" {

as_setup *as_setup_create_ex(const double *vertices, size_t nv, const int64_t *faces, size_t nf,
                             const int64_t *face_material, const double *beta, size_t materials,
                             const double *branches, const int64_t *branch_material, size_t nbranch,
                             const double *origin, double cell, size_t nx, size_t ny, size_t nz,
                             const double *source, const double *probes, size_t nprobes, double edge_admittance,
                             size_t layer, double layer_damping, size_t stencil, const float *speed,
                             const float *flow_x, const float *flow_y, size_t probes_per_listener, int threads,
                             unsigned flags) {
    g_setup_error.clear();
    as_setup *h = new (std::nothrow) as_setup();
    if (h == nullptr) {
        g_setup_error = "out of host memory";
        return nullptr;
    }
    try {
        Setup &s = h->s;
        if (nf == 0) throw std::invalid_argument("the mesh has no faces");
        if (flags & ~static_cast<unsigned>(AS_SETUP_ON_DEVICE | AS_SETUP_FITTED)) throw std::invalid_argument("unknown set-up flags");
        s.on_device = (flags & AS_SETUP_ON_DEVICE) != 0;
        s.fitted_walls = (flags & AS_SETUP_FITTED) != 0;
        s.mesh = fitted::Mesh{vertices, nv, faces, nf};
        if (nx * ny * nz >= (1ull << 40)) throw std::invalid_argument("the grid is too large");
        s.materials = materials;
        s.beta.assign(beta, beta + materials);
        s.face_material.assign(face_material, face_material + nf);
        for (size_t f = 0; f < nf && !s.fitted_walls; ++f) {
            V3 p[3];
            for (int k = 0; k < 3; ++k) {
                const int64_t v = faces[3 * f + static_cast<size_t>(k)];
                if (v < 0 || static_cast<size_t>(v) >= nv) throw std::invalid_argument("a face refers to a missing vertex");
                p[k] = V3{vertices[3 * v], vertices[3 * v + 1], vertices[3 * v + 2]};
            }
            Triangle t;
            t.v0 = p[0];
            t.e1 = p[1] - p[0];
            t.e2 = p[2] - p[0];
            const V3 c = cross(t.e1, t.e2);
            const double len = norm(c);
            if (!(len > 0.0)) throw std::invalid_argument("a face of the mesh has no area");
            t.normal = c * (1.0 / len);
            t.lo = {std::min({p[0].x, p[1].x, p[2].x}), std::min({p[0].y, p[1].y, p[2].y}),
                    std::min({p[0].z, p[1].z, p[2].z})};
            t.hi = {std::max({p[0].x, p[1].x, p[2].x}), std::max({p[0].y, p[1].y, p[2].y}),
                    std::max({p[0].z, p[1].z, p[2].z})};
            t.material = face_material[f];
            s.tri.push_back(t);
        }
        s.origin = V3{origin[0], origin[1], origin[2]};
        s.cell = cell;
        s.nx = nx;
        s.ny = ny;
        s.nz = nz;
        s.sy = nx;
        s.sz = nx * ny;
        s.n = nx * ny * nz;
        s.layer = layer;
        s.layer_damping = layer_damping;
        s.edge_admittance = edge_admittance;
        s.points = stencil;
        s.threads = threads > 0 ? threads : default_threads();
        // A hundred times the reference's tie (kTie, 1e-10 cells); beyond it, and off the
        // faces' edges, its casts and tie rules decide as the plain geometry does.
        s.guard = 1e-8 * cell;
        std::vector<V3> points;
        for (size_t p = 0; p < nprobes; ++p) points.push_back(V3{probes[3 * p], probes[3 * p + 1], probes[3 * p + 2]});
        s.build(branches, branch_material, nbranch, V3{source[0], source[1], source[2]}, points, probes_per_listener,
                speed, flow_x, flow_y);
        return h;
    } catch (const std::exception &e) {
        g_setup_error = e.what();
    }
    delete h;
    return nullptr;
}

as_setup *as_setup_create(const double *vertices, size_t nv, const int64_t *faces, size_t nf,
                          const int64_t *face_material, const double *beta, size_t materials,
                          const double *branches, const int64_t *branch_material, size_t nbranch,
                          const double *origin, double cell, size_t nx, size_t ny, size_t nz,
                          const double *source, const double *probes, size_t nprobes, double edge_admittance,
                          size_t layer, double layer_damping, size_t stencil, const float *speed,
                          size_t probes_per_listener, int threads) {
    return as_setup_create_ex(vertices, nv, faces, nf, face_material, beta, materials, branches, branch_material,
                              nbranch, origin, cell, nx, ny, nz, source, probes, nprobes, edge_admittance, layer,
                              layer_damping, stencil, speed, nullptr, nullptr, probes_per_listener, threads, 0u);
}

const char *as_setup_create_error(void) { return g_setup_error.c_str(); }

int as_setup_layout(const as_setup *h, as_layout *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    const SpecialList &sp = s.specials;
    std::memset(out, 0, sizeof(*out));
    out->nx = s.nx;
    out->ny = s.ny;
    out->nz = s.nz;
    out->stencil = s.points;
    out->weights = s.weights.data();
    out->n_weights = static_cast<uint64_t>(s.st.count);
    out->centre = s.st.centre;
    out->total = s.st.total;
    out->scale = s.scale.empty() ? nullptr : s.scale.data();
    out->layer_scale = s.layer_scale.empty() ? nullptr : s.layer_scale.data();
    out->types = s.type.data();
    out->specials = sp.size();
    out->special = sp.cell.data();
    out->link_start = sp.link_start.data();
    out->links = sp.link_direction.size();
    out->link_direction = sp.link_direction.data();
    out->link_weight = sp.link_weight.data();
    out->coef_a = sp.coef_a.data();
    out->coef_b = sp.coef_b.data();
    out->coef_inv = sp.coef_inv.data();
    out->source = sp.source.data();
    out->state_start = sp.state_start.data();
    out->states = sp.state_branch.size();
    out->state_branch = sp.state_branch.data();
    out->state_weight = sp.state_weight.data();
    out->branches = s.branches.now.size();
    out->branch_now = s.branches.now.data();
    out->branch_before = s.branches.before.data();
    out->branch_gain = s.branches.gain.data();
    out->layer_cells = s.layer_cells.size();
    out->layer_cell = s.layer_cells.data();
    for (int k = 0; k < 6; ++k) {
        out->keep[k] = s.keep[static_cast<size_t>(k)].empty() ? nullptr : s.keep[static_cast<size_t>(k)].data();
        out->push[k] = s.push[static_cast<size_t>(k)].empty() ? nullptr : s.push[static_cast<size_t>(k)].data();
    }
    out->probes = s.probe_start.size() - 1;
    out->probe_start = s.probe_start.data();
    out->taps = s.probe_cells.size();
    out->probe_cells = s.probe_cells.data();
    out->probe_weights = s.probe_weights.data();
    out->flow_x = s.flow_x.empty() ? nullptr : s.flow_x.data();
    out->flow_y = s.flow_y.empty() ? nullptr : s.flow_y.data();
    out->flow_mask = s.flow_mask.empty() ? nullptr : s.flow_mask.data();
    for (size_t d = 0; d < 3; ++d) {
        out->wind[d] = s.flow_wind[d].empty() ? nullptr : s.flow_wind[d].data();
        out->loop[d] = s.flow_loop[d].empty() ? nullptr : s.flow_loop[d].data();
    }
    return AS_OK;
}

int as_setup_info(const as_setup *h, uint64_t *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    out[0] = s.air;
    out[1] = s.boundary_cells;
    out[2] = s.specials.state_branch.size();
    out[3] = s.layer_cells.size();
    out[4] = s.probe_start.size() - 1;
    out[5] = s.open ? 1 : 0;
    uint64_t sides = 0;
    for (size_t k = 0; k < 6; ++k) sides |= (s.open_sides[k] ? 1ull : 0ull) << k;
    out[6] = sides;
    out[7] = s.source_on_face ? 1 : 0;  // fitted walls only: on cell faces it is refused
    return AS_OK;
}

int as_setup_wind(const as_setup *h, int *round, double *fastest, uint64_t *sweeps) {
    if (h == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    if (round) *round = s.flow_round ? 1 : 0;
    if (fastest) *fastest = static_cast<double>(s.flow_fastest);
    if (sweeps) *sweeps = s.flow_sweeps;
    return AS_OK;
}

int as_setup_probe(const as_setup *h, size_t p, int *found, int *in_sight) {
    if (h == nullptr || p >= h->s.probe_found.size()) return AS_EINVAL;
    if (found) *found = h->s.probe_found[p];
    if (in_sight) *in_sight = h->s.probe_in_sight[p];
    return AS_OK;
}

int as_setup_cut_cells(const as_setup *h, uint64_t *rows, const int64_t **cell, const double **air_volume,
                       const double **open_area, const double **wall_area, const double **sliver,
                       const double **special_volume, const uint8_t **kind, double *stats) {
    if (h == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    if (!s.fitted_walls) return AS_EINVAL;
    if (rows) *rows = s.cut_rows.size();
    if (cell) *cell = s.cut_rows.data();
    if (air_volume) *air_volume = s.cut_air_volume.data();
    if (open_area) *open_area = s.cut_open_area.data();
    if (wall_area) *wall_area = s.cut_wall_area.data();
    if (sliver) *sliver = s.cut_sliver.data();
    if (special_volume) *special_volume = s.special_volume.data();
    if (kind) *kind = s.cut_kind.data();
    if (stats) {
        const fitted::FitStats &c = s.cut_stats;
        const double v[] = {static_cast<double>(c.cubes), static_cast<double>(c.regions), static_cast<double>(c.ties),
                            static_cast<double>(c.farthest), static_cast<double>(c.conduit_links),
                            static_cast<double>(c.conduit_scaled), c.far_volume, c.diagonal_area, c.dropped_area,
                            c.edge_area, c.air_volume, c.conduit_least, c.conduit_wall_moved,
                            static_cast<double>(c.faces), static_cast<double>(c.planes), static_cast<double>(c.leaves),
                            static_cast<double>(c.squares), static_cast<double>(c.quantum_bits)};
        std::memcpy(stats, v, sizeof v);
    }
    return AS_OK;
}

void as_setup_destroy(as_setup *h) { delete h; }

}  // extern "C"
