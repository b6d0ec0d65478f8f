// The wave solver's set-up on walls on cell faces: from a triangle mesh, its materials and
// a grid, the layout the time step reads (the reference's WaveSolver constructor,
// cpp/include/acoustic_spaces/fdtd.hpp:360-690, and its layout(), 727-782).
//
// What it computes, in the reference's order and arithmetic:
//  - the stencil's weights, lowered by the stability margin and rounded down
//    (set_up_stencil, fdtd.hpp:1649-1673; lowered/down, 2127-2135);
//  - each branch's coefficients (fdtd.hpp:391-411);
//  - the cut links: bit b of cut[i] when the segment between the centres of cell i and its
//    neighbour along the b-th direction of the + half crosses a face (fdtd.hpp:446-477);
//  - the air: every cell the source's visible cells reach through links that are not
//    blocked (visible_cells, seen_cells 3291-3365; flood 3367-3382);
//  - the matched layer's cells and coefficients (fdtd.hpp:521-542, 2219-2240);
//  - the boundary cells: marks, folded and pair-averaged link weights, per-material
//    boundary weights and branch states (fdtd.hpp:546-625, raw_weights 2162-2190,
//    add_special 3499-3517);
//  - the source's cells, by the cubic reading stencil (reading_cells, 2361-2488), and the
//    probes' cells and weights.
//
// Geometry. The reference decides crossings with a BVH and a set of tie rules for points
// that lie on faces or segments that pass through their edges (kTie, crossing,
// fdtd.hpp:2475-3257). This set-up decides them exactly where no tie can arise: a segment
// crosses a face when its ends lie strictly on either side of the face's plane and the
// point where it meets the plane lies inside the triangle. Every decision is made with a
// margin (kGuard from a plane, kBarycentric from an edge) far wider than the reference's
// tie and edge tolerances (1e-10 cells, 1e-9 barycentric); a decision inside a margin
// that could change the layout raises "degenerate" instead of guessing, so every layout
// this code hands out is the reference's. A meeting near an edge shared with another face
// of the same material at the same point, or farther than a firm crossing, changes nothing
// and is accepted. Sources and probes on faces (the reference's leans) are refused.

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

namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
    double operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};
inline V3 operator+(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
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

inline V3 cross(const V3 &a, const V3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3 &a) { return std::sqrt(dot(a, a)); }

struct Degenerate : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

constexpr uint8_t kSolid = 0, kInterior = 1, kSpecial = 2, kLayer = 3;
constexpr double kMargin = 1.0 / 262144.0;  // 2^-18 (fdtd.hpp:17-27)
constexpr int kReach = 4;                   // cells a side of a cubic reading (fdtd.hpp:2321)
constexpr double kNear = 3.0;               // probes nearer the source read trilinearly
constexpr double kBarycentric = 1e-8;       // the edge margin (see above)
constexpr size_t npos = static_cast<size_t>(-1);

constexpr int kSteps[18][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

int opposite(int k) { return k < 6 ? k ^ 1 : 6 + 4 * ((k - 6) / 4) + 3 - (k - 6) % 4; }

std::array<int, 2> steps_of(int k) {
    static constexpr int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const int p = (k - 6) / 4, sa = (k - 6) % 2, sb = ((k - 6) / 2) % 2;
    return {2 * pairs[p][0] + sa, 2 * pairs[p][1] + sb};
}

// The largest float not above w; a link's weight lowered by the margin (fdtd.hpp:2127-2135).
float down(double w) {
    float f = static_cast<float>(w);
    if (static_cast<double>(f) > w) f = std::nextafter(f, 0.0f);
    return f;
}
float lowered(double w) { return down(w * (1.0 - kMargin)); }

struct Triangle {
    V3 v0, e1, e2, normal;
    V3 lo, hi;  // bounding box
    int64_t material = 0;
};

// A segment's meeting with the faces: the nearest firm crossing and whether any meeting
// lies within the margins.
struct Hit {
    double t = -1.0;
    int64_t face = -1;
};

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
    size_t stencil = 7;
    int directions = 6;
    int threads = 1;
    double guard = 0;  // kGuard: distance from a plane within which an end counts as on it

    // stencil
    std::array<double, 18> exact{}, boundary{};
    std::array<float, 18> weight{};
    std::array<ptrdiff_t, 18> offset{};
    std::array<int, 18> bit{};
    float centre = 0, total = 0;

    // layout
    std::vector<float> scale, layer_scale;
    std::vector<uint8_t> type;
    std::vector<uint16_t> cut;
    std::vector<int64_t> special;
    std::vector<uint64_t> link_start{0};
    std::vector<uint8_t> link_direction;
    std::vector<float> link_weight;
    std::vector<float> coef_a, coef_b, coef_inv, source;
    std::vector<uint64_t> state_start{0};
    std::vector<int32_t> state_branch;
    std::vector<float> state_weight;
    std::vector<float> branch_now, branch_before, branch_gain;
    std::vector<double> branch_a;
    std::vector<int64_t> layer_cells;
    std::array<std::vector<float>, 6> keep, push;
    std::vector<uint64_t> probe_start{0}, probe_cells;
    std::vector<double> probe_weights;
    std::vector<uint8_t> probe_found, probe_in_sight;
    size_t air = 0, boundary_cells = 0;
    bool open = false;
    std::array<bool, 6> open_sides{};

    // triangle bins for crossing queries
    size_t bin = 8, bx = 0, by = 0, bz = 0;
    std::vector<uint32_t> bin_start, bin_faces;
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
    size_t neighbour(size_t i, int k) const { return static_cast<size_t>(static_cast<ptrdiff_t>(i) + offset[k]); }
    bool edge(size_t i) const {
        const size_t x = i % nx, y = (i / nx) % ny, z = i / sz;
        return x == 0 || y == 0 || z == 0 || x + 1 == nx || y + 1 == ny || z + 1 == nz;
    }
    double cell_scale(size_t i) const { return scale.empty() ? 1.0 : static_cast<double>(scale[i]); }
    bool is_cut(size_t i, int k) const {
        if (bit[k] >= 0) return (cut[i] >> bit[k]) & 1u;
        return (cut[neighbour(i, k)] >> bit[opposite(k)]) & 1u;
    }
    bool blocked(size_t i, int k) const { return edge(neighbour(i, k)) || is_cut(i, k); }
    // edge(neighbour(i, k)) for cell i at (x, y, z), without dividing. A step past the
    // grid lands, in neighbour()'s indexing, on a cell of the far edge: on the edge too.
    bool neighbour_on_edge(size_t x, size_t y, size_t z, int k) const {
        const long long to[3] = {static_cast<long long>(x) + kSteps[k][0], static_cast<long long>(y) + kSteps[k][1],
                                 static_cast<long long>(z) + kSteps[k][2]};
        const long long size[3] = {static_cast<long long>(nx), static_cast<long long>(ny), static_cast<long long>(nz)};
        for (int a = 0; a < 3; ++a)
            if (to[a] <= 0 || to[a] >= size[a] - 1) return true;
        return false;
    }
    // blocked(i, k) for cell i at (x, y, z), without dividing.
    bool blocked_at(size_t i, int k, size_t x, size_t y, size_t z) const {
        return neighbour_on_edge(x, y, z, k) || is_cut(i, k);
    }

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

    void set_up_stencil() {
        directions = stencil == 7 ? 6 : 18;
        const double axial = stencil == 7 ? 1.0 / 3.0 : 0.25;
        const double diagonal = stencil == 7 ? 0.0 : 0.125;
        const double courant = std::sqrt(axial + 4.0 * diagonal);
        int bits = 0;
        for (int k = 0; k < 18; ++k) {
            exact[k] = k < directions ? (k < 6 ? axial : diagonal) : 0.0;
            weight[k] = lowered(exact[k]);
            boundary[k] = exact[k] / (2.0 * courant);
            offset[k] = kSteps[k][0] + kSteps[k][1] * static_cast<ptrdiff_t>(sy) + kSteps[k][2] * static_cast<ptrdiff_t>(sz);
            const int *d = kSteps[k];
            const bool plus = d[2] > 0 || (d[2] == 0 && (d[1] > 0 || (d[1] == 0 && d[0] > 0)));
            bit[k] = k < directions && plus ? bits++ : -1;
        }
        double sum = 0.0;
        for (int k = 0; k < directions; ++k) sum += weight[k];
        centre = static_cast<float>(2.0 - sum);
        total = static_cast<float>(sum);
    }

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
                                for (int k = 0; k < directions; ++k) {
                                    if (bit[k] < 0) continue;
                                    const long long tx = static_cast<long long>(x) + kSteps[k][0];
                                    const long long ty = static_cast<long long>(y) + kSteps[k][1];
                                    const long long tz = static_cast<long long>(z) + kSteps[k][2];
                                    if (tx < 0 || ty < 0 || tz < 0 || tx >= static_cast<long long>(nx) ||
                                        ty >= static_cast<long long>(ny) || tz >= static_cast<long long>(nz))
                                        continue;
                                    double frac = 0;
                                    const int r = meet(f, c, centre_at(tx, ty, tz), frac);
                                    if (r == 1) {
                                        cut[i] = static_cast<uint16_t>(cut[i] | (1u << bit[k]));
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
                if (!((cut[i] >> bit[k]) & 1u)) throw Degenerate("degenerate geometry: a link passes through a face's edge");
    }

    // The cells round p (up to eight) whose centres p sees, with trilinear weights that sum
    // to one; only cells keep(i) holds (seen_cells, fdtd.hpp:3305-3365, for p on no face).
    template <typename Keep>
    std::vector<std::pair<size_t, double>> seen_cells(const V3 &p, Keep keep) const {
        std::vector<std::pair<size_t, double>> out;
        const double q[3] = {(p.x - origin.x) / cell - 0.5, (p.y - origin.y) / cell - 0.5, (p.z - origin.z) / cell - 0.5};
        const double size[3] = {static_cast<double>(nx), static_cast<double>(ny), static_cast<double>(nz)};
        long long base[3];
        double frac[3];
        for (int a = 0; a < 3; ++a) {
            if (!(q[a] >= 1.0 && q[a] <= size[a] - 3.0)) return out;
            base[a] = static_cast<long long>(std::floor(q[a]));
            frac[a] = q[a] - static_cast<double>(base[a]);
        }
        std::pair<size_t, double> seen[8];
        int count = 0;
        for (int corner = 0; corner < 8; ++corner) {
            double w = 1.0;
            size_t i = 0, stride = 1;
            for (int a = 0; a < 3; ++a) {
                const int up = (corner >> a) & 1;
                w *= up ? frac[a] : 1.0 - frac[a];
                i += static_cast<size_t>(base[a] + up) * stride;
                stride *= a == 0 ? nx : (a == 1 ? ny : 1);
            }
            if (w <= 0.0) continue;
            if (!keep(i)) continue;
            if (!sees(p, centre_of(i))) continue;
            seen[count++] = {i, w};
        }
        double sum = 0.0;
        for (int k = 0; k < count; ++k) {
            out.push_back(seen[k]);
            sum += seen[k].second;
        }
        if (sum < 1e-9) {
            out.clear();
            return out;
        }
        for (auto &cw : out) cw.second /= sum;
        return out;
    }

    std::vector<std::pair<size_t, double>> visible_cells(const V3 &p, bool air_only) const {
        return seen_cells(p, [&](size_t i) { return !air_only || type[i] != kSolid; });
    }

    // The cubic reading of p (reading_cells, fdtd.hpp:2361-2488), for p on no face.
    std::vector<std::pair<size_t, double>> reading_cells(const V3 &p, bool cubic) const {
        if (!cubic) return visible_cells(p, true);
        const double q[3] = {(p.x - origin.x) / cell - 0.5, (p.y - origin.y) / cell - 0.5, (p.z - origin.z) / cell - 0.5};
        const size_t size[3] = {nx, ny, nz};
        long long below[3], start[3];
        double w3[3][kReach];
        for (int a = 0; a < 3; ++a) {
            below[a] = static_cast<long long>(std::floor(q[a]));
            if (!(q[a] >= 1.0 && q[a] < static_cast<double>(size[a]) - 3.0 && below[a] >= 1)) {
                return visible_cells(p, true);
            }
            start[a] = below[a] - (kReach / 2 - 1);
            const double x = q[a] - static_cast<double>(start[a]);
            for (int j = 0; j < kReach; ++j) {
                double w = 1.0;
                for (int m = 0; m < kReach; ++m) {
                    if (m != j) w *= (x - m) / static_cast<double>(j - m);
                }
                w3[a][j] = w;
            }
        }
        auto index = [&](int x, int y, int z) {
            return static_cast<size_t>(start[0] + x) + static_cast<size_t>(start[1] + y) * sy +
                   static_cast<size_t>(start[2] + z) * sz;
        };
        bool own[kReach][kReach][kReach];
        for (int z = 0; z < kReach; ++z)
            for (int y = 0; y < kReach; ++y)
                for (int x = 0; x < kReach; ++x) {
                    const size_t i = index(x, y, z);
                    bool s = type[i] == kInterior || type[i] == kSpecial;
                    if (s) s = sees(p, centre_of(i));
                    own[z][y][x] = s;
                }
        bool keep_layer[3][kReach];
        for (auto &axis : keep_layer) std::fill(std::begin(axis), std::end(axis), true);
        for (bool changed = true; changed;) {
            changed = false;
            for (int a = 0; a < 3; ++a) {
                const int b = (a + 1) % 3, c = (a + 2) % 3;
                for (int j = 0; j < kReach; ++j) {
                    if (!keep_layer[a][j]) continue;
                    bool any = false;
                    int at[3];
                    at[a] = j;
                    for (at[b] = 0; at[b] < kReach && !any; ++at[b]) {
                        for (at[c] = 0; at[c] < kReach && !any; ++at[c]) {
                            any = keep_layer[b][at[b]] && keep_layer[c][at[c]] && own[at[2]][at[1]][at[0]];
                        }
                    }
                    if (!any) {
                        keep_layer[a][j] = false;
                        changed = true;
                    }
                }
            }
        }
        int from[3], count[3];
        for (int a = 0; a < 3; ++a) {
            const int inner = kReach / 2 - 1;
            if (keep_layer[a][0] && keep_layer[a][1] && keep_layer[a][2] && keep_layer[a][3]) {
                from[a] = 0;
                count[a] = kReach;
            } else if (keep_layer[a][inner] && keep_layer[a][inner + 1]) {
                from[a] = inner;
                count[a] = 2;
                const double t = q[a] - static_cast<double>(below[a]);
                w3[a][inner] = 1.0 - t;
                w3[a][inner + 1] = t;
            } else if (keep_layer[a][inner] || keep_layer[a][inner + 1]) {
                from[a] = keep_layer[a][inner] ? inner : inner + 1;
                count[a] = 1;
                w3[a][from[a]] = 1.0;
            } else {
                return visible_cells(p, true);
            }
        }
        std::vector<std::pair<size_t, double>> out;
        for (int z = from[2]; z < from[2] + count[2]; ++z)
            for (int y = from[1]; y < from[1] + count[1]; ++y)
                for (int x = from[0]; x < from[0] + count[0]; ++x) {
                    const size_t i = index(x, y, z);
                    if (!own[z][y][x] || (x + 1 < from[0] + count[0] && is_cut(i, 1)) ||
                        (y + 1 < from[1] + count[1] && is_cut(i, 3)) || (z + 1 < from[2] + count[2] && is_cut(i, 5))) {
                        return visible_cells(p, true);
                    }
                    const double w = w3[0][x] * w3[1][y] * w3[2][z];
                    if (w != 0.0) out.emplace_back(i, w);
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
            for (int k = 2; k < directions; ++k) {
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

    void raw_weights(size_t i, double *w) const {
        for (int k = 0; k < 18; ++k) w[k] = 0.0;
        for (int k = 0; k < directions; ++k) {
            if (!blocked(i, k)) {
                w[k] += exact[k];
                continue;
            }
            if (k < 6 || edge(neighbour(i, k))) continue;
            const auto [ka, kb] = steps_of(k);
            const bool behind_a = blocked(i, ka), behind_b = blocked(i, kb);
            if (behind_a && behind_b) continue;
            if (behind_a) {
                w[kb] += exact[k];
            } else if (behind_b) {
                w[ka] += exact[k];
            } else {
                w[ka] += 0.5 * exact[k];
                w[kb] += 0.5 * exact[k];
            }
        }
    }

    void layer_coefficients(int a) {
        const size_t size = a == 0 ? nx : (a == 1 ? ny : nz);
        const double lay = static_cast<double>(layer);
        const double courant = std::sqrt(static_cast<double>(weight[0]));
        auto damping = [&](double from_edge) {
            const double into = std::max(0.0, lay + 0.5 - from_edge) / lay;
            return layer_damping * into * into;
        };
        auto &ck = keep[2 * a], &fk = keep[2 * a + 1];
        auto &cp = push[2 * a], &fp = push[2 * a + 1];
        for (size_t x = 0; x <= size; ++x) {
            const double c = static_cast<double>(x), last = static_cast<double>(size - 1);
            const double dc = damping(std::min(c, last - c));
            const double df = damping(std::min(c - 0.5, last - c + 0.5));
            ck.push_back(static_cast<float>((1.0 - dc) / (1.0 + dc)));
            cp.push_back(static_cast<float>(courant / (1.0 + dc)));
            fk.push_back(static_cast<float>((1.0 - df) / (1.0 + df)));
            fp.push_back(static_cast<float>(courant / (1.0 + df)));
        }
    }

    // The specials of a slab of planes, as the layout's arrays (starts from 0).
    struct Specials {
        std::vector<int64_t> special;
        std::vector<uint64_t> link_start{0};
        std::vector<uint8_t> link_direction;
        std::vector<float> link_weight;
        std::vector<float> coef_a, coef_b, coef_inv, source;
        std::vector<uint64_t> state_start{0};
        std::vector<int32_t> state_branch;
        std::vector<float> state_weight;
        size_t air = 0;
        bool open = false;
        std::array<bool, 6> sides{};
    };

    // Special i's links (the weights w it has, lowered), coefficients and the start of its
    // states (already added) into `out`, whose layout arrays it extends.
    template <typename Out>
    void put_special(Out &out, size_t i, const double *w, double g) const {
        out.special.push_back(static_cast<int64_t>(i));
        const double s = cell_scale(i), r = std::sqrt(s);
        double sum = 0.0;
        for (int k = 0; k < directions; ++k) {
            if (!(w[k] > 0.0)) continue;
            sum += w[k];
            out.link_direction.push_back(static_cast<uint8_t>(k));
            out.link_weight.push_back(lowered(s * w[k]));
        }
        out.link_start.push_back(out.link_direction.size());
        out.coef_a.push_back(static_cast<float>(2.0 - s * sum * (1.0 - kMargin)));
        out.coef_b.push_back(static_cast<float>(1.0 - r * g));
        out.coef_inv.push_back(static_cast<float>(1.0 / (1.0 + r * g)));
        out.source.push_back(0.0f);
        out.state_start.push_back(out.state_branch.size());
    }

    size_t add_special(size_t i, const double *w, double g) {
        type[i] = kSpecial;
        put_special(*this, i, w, g);
        return special.size() - 1;
    }

    size_t special_index(size_t i) const {
        if (type[i] != kSpecial) return npos;
        const auto it = std::find(special.begin(), special.end(), static_cast<int64_t>(i));
        return it == special.end() ? npos : static_cast<size_t>(it - special.begin());
    }

    void build(const double *branches, const int64_t *branch_material, size_t branch_count, const V3 &src,
               const std::vector<V3> &probes, size_t per_listener, const float *speed) {
        if (stencil != 7 && stencil != 19) throw std::invalid_argument("the stencil has 7 or 19 points");
        if (per_listener == 0 || probes.size() % per_listener != 0) {
            throw std::invalid_argument("the probes come in whole groups, one per listener");
        }
        if (layer > 0 && stencil != 7) throw std::invalid_argument("the matched layer needs the 7-point stencil");
        if (nx < 3 || ny < 3 || nz < 3) throw std::invalid_argument("the grid is too small");
        if (!(cell > 0.0)) throw std::invalid_argument("cell must be positive");
        if (2 * layer + 3 > std::min({nx, ny, nz})) throw std::invalid_argument("the matched layer leaves no grid inside it");
        if (!(layer_damping >= 0.0)) throw std::invalid_argument("negative layer damping");
        for (size_t f = 0; f < face_material.size(); ++f) {
            if (face_material[f] < 0 || static_cast<size_t>(face_material[f]) >= materials) {
                throw std::invalid_argument("a face has no material");
            }
        }
        std::vector<std::vector<int32_t>> material_branches(materials);
        for (size_t m = 0; m < materials; ++m) {
            if (!(beta[m] >= 0.0)) throw std::invalid_argument("negative admittance");
        }
        for (size_t j = 0; j < branch_count; ++j) {
            const double m = branches[3 * j], r = branches[3 * j + 1], q = branches[3 * j + 2];
            const double a = m + 0.5 * r + 0.5 * q;
            if (!(m >= 0.0 && r >= 0.0 && q >= 0.0 && a > 0.0)) {
                throw std::invalid_argument("a branch needs mass, resistance and stiffness >= 0");
            }
            if (branch_material[j] < 0 || static_cast<size_t>(branch_material[j]) >= materials) {
                throw std::invalid_argument("a branch has no material");
            }
            material_branches[static_cast<size_t>(branch_material[j])].push_back(static_cast<int32_t>(j));
            branch_now.push_back(static_cast<float>(2.0 * m / a));
            branch_before.push_back(static_cast<float>((2.0 * m + q) / a));
            branch_gain.push_back(static_cast<float>(0.5 / a));
            branch_a.push_back(a);
        }
        set_up_stencil();
        if (speed != nullptr) {
            scale.resize(n);
            for (size_t i = 0; i < n; ++i) {
                const double r = static_cast<double>(speed[i]);
                if (!(r > 0.0 && r <= 1.0)) {
                    throw std::invalid_argument("each cell's speed of sound over the grid's lies in (0, 1]");
                }
                scale[i] = static_cast<float>(r * r);
            }
        }
        SetupClock clock;
        make_bins();
        clock.mark("stencil, bins");
        require_off_faces(src, "the source");
        for (size_t p = 0; p < probes.size(); ++p) require_off_faces(probes[p], "a probe");
        cut_links();
        clock.mark("cut links");

        // Air: the cells the source sees, and all they reach.
        type.assign(n, kSolid);
        std::vector<size_t> queue;
        for (const auto &[i, w] : visible_cells(src, false)) {
            (void)w;
            if (type[i] == kSolid) {
                type[i] = kInterior;
                queue.push_back(i);
            }
        }
        if (queue.empty()) throw std::invalid_argument("the source sees no cell of the wave grid");
        flood(queue);
        clock.mark("air");

        // The matched layer's air, slab by slab of planes on the threads (each cell's class
        // its own; the cells listed in order).
        if (layer > 0) {
            const size_t T = slabs();
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
                            const size_t at[3] = {x, y, z};
                            for (int k = 0; k < 6; ++k) {
                                if (!neighbour_on_edge(x, y, z, k) && is_cut(i, k)) {
                                    throw std::invalid_argument("the model reaches into the matched layer");
                                }
                            }
                            type[i] = kLayer;
                            found[t].push_back(static_cast<int64_t>(i));
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
                for (int k = 0; k < 6; ++k) open_sides[static_cast<size_t>(k)] = open_sides[static_cast<size_t>(k)] || sides[t][static_cast<size_t>(k)];
            }
        }
        if (!layer_cells.empty()) {
            for (int a = 0; a < 3; ++a) layer_coefficients(a);
            for (size_t c = 0; c < layer_cells.size() && !scale.empty(); ++c) {
                layer_scale.push_back(scale[static_cast<size_t>(layer_cells[c])]);
            }
        }

        clock.mark("layer");
        // Boundary cells: marks 1 and 2, then weights, materials and states. Each pass runs
        // on the threads by slabs of planes; the second writes its marks apart (a cell can
        // take mark 2 from several neighbours), and the specials of each slab are joined in
        // the order of the slabs, so of the cells, as one pass in order would list them.
        const size_t T = slabs();
        std::vector<uint8_t> mark(n, 0);
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            for (size_t z = z0; z < z1; ++z)
                for (size_t y = 0; y < ny; ++y)
                    for (size_t x = 0; x < nx; ++x) {
                        const size_t i = x + y * sy + z * sz;
                        if (type[i] != kInterior) continue;
                        for (int k = 0; k < directions && !mark[i]; ++k) mark[i] = blocked_at(i, k, x, y, z);
                    }
        });
        std::vector<uint8_t> second(n, 0);
        in_slabs(T, [&](size_t, size_t z0, size_t z1) {
            std::array<double, 18> wi{};
            for (size_t i = z0 * sz; i < z1 * sz; ++i) {
                if (mark[i] != 1) continue;
                raw_weights(i, wi.data());
                for (int k = 0; k < directions; ++k) {
                    const size_t j = neighbour(i, k);
                    if (wi[k] != exact[k] && type[j] == kInterior && !mark[j]) {
                        __atomic_store_n(&second[j], static_cast<uint8_t>(1), __ATOMIC_RELAXED);
                    }
                }
            }
        });
        for (size_t i = 0; i < n; ++i)
            if (second[i]) mark[i] = 2;
        std::vector<Specials> parts(T);
        in_slabs(T, [&](size_t t, size_t z0, size_t z1) {
          Specials &out = parts[t];
          Stamps stamps;
          std::array<double, 18> wi{}, wj{};
          for (size_t i = z0 * sz; i < z1 * sz; ++i) {
            if (type[i] != kInterior) continue;
            ++out.air;
            if (!mark[i]) continue;
            raw_weights(i, wi.data());
            std::array<double, 18> w{};
            double beta_sum = 0.0;
            std::array<int32_t, 18> crossed{};
            std::array<double, 18> crossed_weight{};
            size_t kinds = 0;
            auto meet_face = [&](int k, double bw) {
                const size_t j = neighbour(i, k);
                if (edge(j)) {
                    beta_sum += bw * edge_admittance;
                    out.open = true;
                    const size_t at[3] = {j % nx, (j / nx) % ny, j / sz};
                    const size_t size[3] = {nx, ny, nz};
                    for (size_t a = 0; a < 3; ++a) {
                        if (at[a] == 0) out.sides[2 * a] = true;
                        if (at[a] + 1 >= size[a]) out.sides[2 * a + 1] = true;
                    }
                    return;
                }
                int64_t face = crossing(centre_of(i), centre_of(j), stamps);
                if (face < 0) face = crossing(centre_of(j), centre_of(i), stamps);
                if (face < 0) return;
                const auto m = static_cast<int32_t>(face_material[static_cast<size_t>(face)]);
                beta_sum += bw * beta[static_cast<size_t>(m)];
                size_t kind = 0;
                while (kind < kinds && crossed[kind] != m) ++kind;
                if (kind == kinds) crossed[kinds++] = m;
                crossed_weight[kind] += bw;
            };
            for (int k = 0; k < directions; ++k) {
                const size_t j = neighbour(i, k);
                if (!edge(j) && type[j] != kSolid && type[j] != kLayer && mark[j]) {
                    raw_weights(j, wj.data());
                    w[k] = 0.5 * (wi[k] + wj[opposite(k)]);
                } else if (!edge(j) && type[j] != kSolid) {
                    w[k] = 0.5 * (wi[k] + exact[k]);
                }
                if (!blocked(i, k)) continue;
                if (k < 6) {
                    meet_face(k, boundary[k]);
                    continue;
                }
                const auto [ka, kb] = steps_of(k);
                const bool behind_a = blocked(i, ka), behind_b = blocked(i, kb);
                if (behind_a) meet_face(ka, boundary[k]);
                if (behind_b) meet_face(kb, boundary[k]);
                if (!behind_a && !behind_b) meet_face(k, boundary[k]);
            }
            double g = beta_sum;
            const double r = std::sqrt(cell_scale(i));
            for (size_t kind = 0; kind < kinds; ++kind) {
                for (const int32_t jb : material_branches[static_cast<size_t>(crossed[kind])]) {
                    g += crossed_weight[kind] * (0.5 / branch_a[static_cast<size_t>(jb)]);
                    out.state_branch.push_back(jb);
                    out.state_weight.push_back(static_cast<float>(r * crossed_weight[kind]));
                }
            }
            put_special(out, i, w.data(), g);
          }
        });
        for (const Specials &part : parts) {
            const uint64_t links0 = link_direction.size(), states0 = state_branch.size();
            for (const int64_t i : part.special) type[static_cast<size_t>(i)] = kSpecial;
            special.insert(special.end(), part.special.begin(), part.special.end());
            for (size_t q = 1; q < part.link_start.size(); ++q) link_start.push_back(links0 + part.link_start[q]);
            for (size_t q = 1; q < part.state_start.size(); ++q) state_start.push_back(states0 + part.state_start[q]);
            link_direction.insert(link_direction.end(), part.link_direction.begin(), part.link_direction.end());
            link_weight.insert(link_weight.end(), part.link_weight.begin(), part.link_weight.end());
            coef_a.insert(coef_a.end(), part.coef_a.begin(), part.coef_a.end());
            coef_b.insert(coef_b.end(), part.coef_b.begin(), part.coef_b.end());
            coef_inv.insert(coef_inv.end(), part.coef_inv.begin(), part.coef_inv.end());
            source.insert(source.end(), part.source.begin(), part.source.end());
            state_branch.insert(state_branch.end(), part.state_branch.begin(), part.state_branch.end());
            state_weight.insert(state_weight.end(), part.state_weight.begin(), part.state_weight.end());
            air += part.air;
            open = open || part.open;
            for (int k = 0; k < 6; ++k) open_sides[static_cast<size_t>(k)] = open_sides[static_cast<size_t>(k)] || part.sides[static_cast<size_t>(k)];
        }
        boundary_cells = special.size();
        clock.mark("marks, specials");

        // The source's cells.
        for (const auto &[i, w] : visible_cells(src, true)) {
            (void)w;
            if (type[i] == kLayer) throw std::invalid_argument("the source is in the matched layer");
        }
        for (const auto &[i, w] : reading_cells(src, true)) {
            size_t k = special_index(i);
            if (k == npos) k = add_special(i, exact.data(), 0.0);
            source[k] += static_cast<float>(cell_scale(i) * w);
        }
        // The probes' cells.
        for (size_t k = 0; k < probes.size(); ++k) {
            const V3 &p = probes[k];
            const V3 d = p - src;
            const bool near = norm(d) < (kNear - 1e-9) * cell;
            const auto cells = reading_cells(p, !near);
            for (const auto &[i, w] : cells) {
                probe_cells.push_back(i);
                probe_weights.push_back(w);
            }
            probe_start.push_back(probe_cells.size());
            probe_found.push_back(cells.empty() ? 0 : 1);
        }
        for (size_t k = 0; k < probes.size(); ++k) {
            const size_t l = k - k % per_listener;
            probe_in_sight.push_back(k == l || sees(probes[l], probes[k]) ? 1 : 0);
        }
        clock.mark("source, probes");
    }
};

thread_local std::string g_setup_error;

}  // namespace

struct as_setup {
    Setup s;
};

extern "C" {

as_setup *as_setup_create(const double *vertices, size_t nv, const int64_t *faces, size_t nf,
                          const int64_t *face_material, const double *beta, size_t materials,
                          const double *branches, const int64_t *branch_material, size_t nbranch,
                          const double *origin, double cell, size_t nx, size_t ny, size_t nz,
                          const double *source, const double *probes, size_t nprobes, double edge_admittance,
                          size_t layer, double layer_damping, size_t stencil, const float *speed,
                          size_t probes_per_listener, int threads) {
    g_setup_error.clear();
    as_setup *h = new (std::nothrow) as_setup();
    if (h == nullptr) {
        g_setup_error = "out of host memory";
        return nullptr;
    }
    try {
        Setup &s = h->s;
        if (nf == 0) throw std::invalid_argument("the mesh has no faces");
        if (nx * ny * nz >= (1ull << 40)) throw std::invalid_argument("the grid is too large");
        s.materials = materials;
        s.beta.assign(beta, beta + materials);
        s.face_material.assign(face_material, face_material + nf);
        V3 lo{std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
              std::numeric_limits<double>::infinity()};
        V3 hi{-lo.x, -lo.y, -lo.z};
        for (size_t f = 0; f < nf; ++f) {
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
            t.normal = len > 0.0 ? c * (1.0 / len) : V3{};
            t.lo = {std::min({p[0].x, p[1].x, p[2].x}), std::min({p[0].y, p[1].y, p[2].y}),
                    std::min({p[0].z, p[1].z, p[2].z})};
            t.hi = {std::max({p[0].x, p[1].x, p[2].x}), std::max({p[0].y, p[1].y, p[2].y}),
                    std::max({p[0].z, p[1].z, p[2].z})};
            t.material = face_material[f];
            lo = {std::min(lo.x, t.lo.x), std::min(lo.y, t.lo.y), std::min(lo.z, t.lo.z)};
            hi = {std::max(hi.x, t.hi.x), std::max(hi.y, t.hi.y), std::max(hi.z, t.hi.z)};
            if (!(len > 0.0)) throw std::invalid_argument("a face of the mesh has no area");
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
        s.stencil = stencil;
        s.threads = threads > 0 ? threads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
        const double scale = std::max(norm(hi - lo), 1e-9);
        // A hundred times the reference's tie (kTie, 1e-10 cells); beyond it, and off the
        // faces' edges, its casts and tie rules decide as the plain geometry does.
        (void)scale;
        s.guard = 1e-8 * cell;
        std::vector<V3> points;
        for (size_t p = 0; p < nprobes; ++p) points.push_back(V3{probes[3 * p], probes[3 * p + 1], probes[3 * p + 2]});
        s.build(branches, branch_material, nbranch, V3{source[0], source[1], source[2]}, points, probes_per_listener,
                speed);
        return h;
    } catch (const std::exception &e) {
        g_setup_error = e.what();
    }
    delete h;
    return nullptr;
}

const char *as_setup_create_error(void) { return g_setup_error.c_str(); }

int as_setup_layout(const as_setup *h, as_layout *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    std::memset(out, 0, sizeof(*out));
    out->nx = s.nx;
    out->ny = s.ny;
    out->nz = s.nz;
    out->stencil = s.stencil;
    out->weights = s.weight.data();
    out->n_weights = static_cast<uint64_t>(s.directions);
    out->centre = s.centre;
    out->total = s.total;
    out->scale = s.scale.empty() ? nullptr : s.scale.data();
    out->layer_scale = s.layer_scale.empty() ? nullptr : s.layer_scale.data();
    out->types = s.type.data();
    out->specials = s.special.size();
    out->special = s.special.data();
    out->link_start = s.link_start.data();
    out->links = s.link_direction.size();
    out->link_direction = s.link_direction.data();
    out->link_weight = s.link_weight.data();
    out->coef_a = s.coef_a.data();
    out->coef_b = s.coef_b.data();
    out->coef_inv = s.coef_inv.data();
    out->source = s.source.data();
    out->state_start = s.state_start.data();
    out->states = s.state_branch.size();
    out->state_branch = s.state_branch.data();
    out->state_weight = s.state_weight.data();
    out->branches = s.branch_now.size();
    out->branch_now = s.branch_now.data();
    out->branch_before = s.branch_before.data();
    out->branch_gain = s.branch_gain.data();
    out->layer_cells = s.layer_cells.size();
    out->layer_cell = s.layer_cells.data();
    for (int k = 0; k < 6; ++k) {
        out->keep[k] = s.keep[k].empty() ? nullptr : s.keep[k].data();
        out->push[k] = s.push[k].empty() ? nullptr : s.push[k].data();
    }
    out->probes = s.probe_start.size() - 1;
    out->probe_start = s.probe_start.data();
    out->taps = s.probe_cells.size();
    out->probe_cells = s.probe_cells.data();
    out->probe_weights = s.probe_weights.data();
    return AS_OK;
}

int as_setup_info(const as_setup *h, uint64_t *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    const Setup &s = h->s;
    out[0] = s.air;
    out[1] = s.boundary_cells;
    out[2] = s.state_branch.size();
    out[3] = s.layer_cells.size();
    out[4] = s.probe_start.size() - 1;
    out[5] = s.open ? 1 : 0;
    uint64_t sides = 0;
    for (int k = 0; k < 6; ++k) sides |= (s.open_sides[static_cast<size_t>(k)] ? 1ull : 0ull) << k;
    out[6] = sides;
    out[7] = 0;  // the source on a face: refused here
    return AS_OK;
}

int as_setup_probe(const as_setup *h, size_t p, int *found, int *in_sight) {
    if (h == nullptr || p >= h->s.probe_found.size()) return AS_EINVAL;
    if (found) *found = h->s.probe_found[p];
    if (in_sight) *in_sight = h->s.probe_in_sight[p];
    return AS_OK;
}

void as_setup_destroy(as_setup *h) { delete h; }

}  // extern "C"
