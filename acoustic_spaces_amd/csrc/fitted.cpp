// Fitted walls: each air cell's share of the air, open faces and walls, from the cut of a
// closed triangle mesh by the grid's cells. It stands in for the reference's CutCells
// (cpp/include/acoustic_spaces/cut_cells.hpp:1-147 states its rules) on meshes whose faces
// close inside the grid and lie off the grid's planes and corners.
//
// Design.
//  - The lattice. Coordinates are taken once onto 64-bit integers (the reference's frame,
//    cut_cells.hpp:11-16, so that each cell's width on it, and so each uncut cell's volume,
//    is the reference's to the bit). Mesh vertices with one lattice point are one vertex.
//  - Planes. A face's plane is an integer normal (primitive, first nonzero component
//    positive) and offset, exact in 128 and 256 bits, so coplanar faces share one plane.
//  - A cell that a face meets with positive area is a cube. In each cube the geometry runs
//    in long double in the cube's own frame (cells from its lattice corner), where every
//    mesh vertex's coordinates are exact. The cube is split by the planes of its faces, in
//    the planes' order, a piece (a leaf) being split by a plane only where a face on that
//    plane has positive area inside it. Leaves are convex: a list of half-spaces.
//  - Incidence by construction. Every point of a polygon carries the planes it was made on
//    and, where it is one, its mesh vertex. A point is on a plane when the plane is one of
//    its own, when its mesh vertex is one of the plane's, or when two of its own planes meet
//    in a line through two mesh vertices the plane holds (a welded edge). Otherwise its side
//    is the sign of its distance, which must clear 1e-12 cells or the input is refused. So a
//    welded edge, the corner of three walls and a face's edge on a neighbour's plane are met
//    exactly, and a polygon keeps positive area after a cut exactly when one of its points
//    lies strictly on the kept side.
//  - Regions. Two leaves on either side of a plane join where their facets overlap in a
//    part of positive area that no face on the plane covers; the covered parts are walls of
//    both, by the covering face's material. Squares (the faces between a cube and its six
//    neighbours) are cut by the leaves of the cells on both sides; each part of positive
//    area joins the regions on its two sides.
//  - The air, the owners, the rows, the sliver thickness and the conduit rule follow the
//    reference's rules (cut_cells.hpp:66-101; air 4410-4640, owners 4641-4720, sum_up
//    4846-4970, thin_conduits 5185-5352), on this file's own data.
//
// Measures. Volumes by the divergence theorem about the cube's centre, areas by the
// shoelace sum of a polygon projected onto its plane's largest axis. They are the
// reference's measures to within rounding (about 1e-18 of a cell in long double), not to
// the bit; the decisions (cut or not, open or not, which region holds a point) are exact.
//
// Not here (refused): faces lying on a plane of the grid's cells (the reference's
// membranes), open edges of the mesh inside the grid (its refusal (i) and its slits), and a
// cell corner, edge or face on a face's plane within 1e-12 of a cell where no construction
// places it there.

#include "fitted.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace fitted {
namespace {

using i128 = __int128;
using u128 = unsigned __int128;
using real = long double;

// ---- wide integers ------------------------------------------------------------------------

// A signed 256-bit integer, two's complement in four words.
struct I256 {
    uint64_t w[4] = {0, 0, 0, 0};
};

I256 operator+(const I256 &a, const I256 &b) {
    I256 r;
    u128 c = 0;
    for (int k = 0; k < 4; ++k) {
        c += static_cast<u128>(a.w[k]) + b.w[k];
        r.w[k] = static_cast<uint64_t>(c);
        c >>= 64;
    }
    return r;
}

I256 operator-(const I256 &a) {
    I256 r;
    for (int k = 0; k < 4; ++k) r.w[k] = ~a.w[k];
    I256 one;
    one.w[0] = 1;
    return r + one;
}

I256 operator-(const I256 &a, const I256 &b) { return a + (-b); }

bool negative(const I256 &a) { return (a.w[3] >> 63) != 0; }
bool is_zero(const I256 &a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0; }
int sign(const I256 &a) { return negative(a) ? -1 : (is_zero(a) ? 0 : 1); }
bool operator==(const I256 &a, const I256 &b) { return std::memcmp(a.w, b.w, sizeof a.w) == 0; }
bool operator<(const I256 &a, const I256 &b) { return negative(a - b); }

// a b, exactly (|a| < 2^127, any b).
I256 mul(i128 a, int64_t b) {
    const bool flip = (a < 0) != (b < 0);
    const u128 ua = a < 0 ? static_cast<u128>(0) - static_cast<u128>(a) : static_cast<u128>(a);
    const uint64_t ub = b < 0 ? uint64_t{0} - static_cast<uint64_t>(b) : static_cast<uint64_t>(b);
    const u128 p0 = static_cast<u128>(static_cast<uint64_t>(ua)) * ub;
    const u128 p1 = static_cast<u128>(static_cast<uint64_t>(ua >> 64)) * ub;
    I256 r;
    r.w[0] = static_cast<uint64_t>(p0);
    const u128 mid = (p0 >> 64) + static_cast<uint64_t>(p1);
    r.w[1] = static_cast<uint64_t>(mid);
    const u128 hi = (mid >> 64) + (p1 >> 64);
    r.w[2] = static_cast<uint64_t>(hi);
    r.w[3] = static_cast<uint64_t>(hi >> 64);
    return flip ? -r : r;
}

real to_real(I256 a) {
    const bool neg = negative(a);
    if (neg) a = -a;
    real v = 0;
    for (int k = 3; k >= 0; --k) v = v * 18446744073709551616.0L + static_cast<real>(a.w[k]);
    return neg ? -v : v;
}

u128 gcd(u128 a, u128 b) {
    while (b != 0) {
        const u128 t = a % b;
        a = b;
        b = t;
    }
    return a;
}

u128 mag(i128 x) { return x < 0 ? static_cast<u128>(0) - static_cast<u128>(x) : static_cast<u128>(x); }

// n / d correctly rounded, for 0 <= n <= d < 2^63.
double quotient(int64_t n, int64_t d) {
    if (d < (int64_t{1} << 53)) return static_cast<double>(n) / static_cast<double>(d);
    if (n == 0) return 0.0;
    if (n == d) return 1.0;
    // n / d in [2^-63, 1): 66 bits of the quotient, then round half even on the rest.
    int shift = 0;
    u128 num = static_cast<u128>(n);
    while (num < static_cast<u128>(d)) {
        num <<= 1;
        ++shift;
    }
    // now d <= num < 2 d: the quotient num / d in [1, 2); take 64 more bits.
    num <<= 64;
    const u128 q = num / static_cast<u128>(d);  // in [2^64, 2^65)
    const bool sticky = num % static_cast<u128>(d) != 0;
    // keep 53 bits: q has 65 significant bits
    const int drop = 12;
    u128 keep = q >> drop;
    const u128 rest = q & ((static_cast<u128>(1) << drop) - 1);
    const u128 half = static_cast<u128>(1) << (drop - 1);
    if (rest > half || (rest == half && (sticky || (keep & 1)))) ++keep;
    return std::ldexp(static_cast<double>(static_cast<uint64_t>(keep)), drop - 64 - shift);
}

struct Refused : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

[[noreturn]] void too_close(const char *what) {
    throw Refused(std::string("fitted walls: ") + what +
                  " within 1e-12 of a cell where the model's own faces do not put it there; this set-up "
                  "decides such meetings only where they follow from the mesh (welded edges and vertices). "
                  "Move the model or the grid by a small fraction of a cell.");
}

// ---- plane ids -------------------------------------------------------------------------------

// A plane of a cube's geometry: a face plane (its index), a plane of the grid's cells
// (axis a at lattice index i), or a face's coverage plane along its edge k (through the
// edge, perpendicular to the face, the face on its + side).
constexpr int64_t kAxisPlane = int64_t{1} << 40, kCoverPlane = int64_t{2} << 40;
int64_t axis_id(int a, size_t i) { return kAxisPlane + 3 * static_cast<int64_t>(i) + a; }
int64_t cover_id(size_t f, int k) { return kCoverPlane + 3 * static_cast<int64_t>(f) + k; }

// A plane in a cube's frame: un . u = off, u in cells from the frame's corner.
struct LP {
    int64_t id = 0;
    real un[3] = {0, 0, 0};
    real off = 0;
};

// A point of a polygon in a cube's frame: its coordinates, its mesh vertex (or -1) and the
// planes it was made on.
struct Vx {
    real u[3] = {0, 0, 0};
    int vid = -1;
    int n = 0;
    int64_t on[6] = {0, 0, 0, 0, 0, 0};
    void add(int64_t id) {
        for (int k = 0; k < n; ++k)
            if (on[k] == id) return;
        if (n < 6) on[n++] = id;
    }
    bool has(int64_t id) const {
        for (int k = 0; k < n; ++k)
            if (on[k] == id) return true;
        return false;
    }
};
using Poly = std::vector<Vx>;

struct Frame {
    int64_t c[3] = {0, 0, 0};
};

// A leaf of a cube: the half-spaces (face plane, side) that cut it out of the cube.
struct Cut {
    int plane;
    int side;
};
using Leaf = std::vector<Cut>;

struct Wall {
    int material;  // -1: the grid's edge
    double area;
};

constexpr int64_t kNone = -1, kEdge = -2;
constexpr double kTol = 1e-12;

// Union-find with path halving.
struct UF {
    std::vector<size_t> up;
    explicit UF(size_t n) : up(n) {
        for (size_t i = 0; i < n; ++i) up[i] = i;
    }
    size_t find(size_t x) {
        while (up[x] != x) {
            up[x] = up[up[x]];
            x = up[x];
        }
        return x;
    }
    void join(size_t a, size_t b) {
        a = find(a);
        b = find(b);
        if (a == b) return;
        if (a < b) std::swap(a, b);
        up[a] = b;  // the lower index is the root
    }
};

template <typename Work>
void parallel(size_t count, int threads, Work work) {
    const size_t T = std::max<size_t>(1, std::min<size_t>(static_cast<size_t>(std::max(threads, 1)), count));
    std::vector<std::exception_ptr> errors(T);
    auto run = [&](size_t t) {
        try {
            work(count * t / T, count * (t + 1) / T);
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

constexpr int kSteps[18][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

int opposite_of(int k) {
    for (int j = 0; j < 18; ++j)
        if (kSteps[j][0] == -kSteps[k][0] && kSteps[j][1] == -kSteps[k][1] && kSteps[j][2] == -kSteps[k][2]) return j;
    return -1;
}

}  // namespace

// ---- the fit --------------------------------------------------------------------------------

struct Fit::Impl {
    size_t n[3] = {0, 0, 0}, sy = 0, sz = 0, N = 0, layer = 0, materials = 0;
    double h = 0;
    int threads = 1;

    // The lattice (cut_cells.hpp:11-16): x -> round((x - X0) 2^s), ties to even.
    int s = 0;
    double S = 0;
    i128 X0s[3] = {0, 0, 0};  // X0 2^s
    std::array<std::vector<int64_t>, 3> G, C;

    // The mesh on the lattice.
    std::vector<std::array<int64_t, 3>> wq;  // welded vertices
    struct FaceX {
        int v[3];
        int plane = -1;
        int material = 0;
        int64_t index = 0;  // the mesh's face index
        int64_t lo[3], hi[3];
        double cover_normal[3][3];  // each edge's coverage plane normal (unit, double)
    };
    std::vector<FaceX> faces;
    struct PlaneX {
        i128 n[3];
        I256 d;
        double unit[3];
        std::vector<int> verts;  // mesh vertices on the plane (exact), ascending
    };
    std::vector<PlaneX> planes;
    std::map<std::pair<int, int>, std::vector<int>> edges;  // welded edge -> faces
    std::map<std::pair<int, size_t>, std::vector<int>> axis_verts;  // (axis, lattice index) -> vertices on it

    // Cubes.
    struct Cube {
        size_t cell = 0;
        std::vector<int> faces;   // faces meeting it with positive area
        std::vector<int> planes;  // their planes, ascending
        std::vector<Leaf> leaves;
        std::vector<int> leaf_region;
        std::vector<double> volume;           // per region
        std::vector<std::vector<Wall>> wall;  // per region
        int centre = 0;                       // the region holding the cell's centre
        size_t leaf_count = 0;
    };
    std::vector<Cube> cubes;
    std::vector<size_t> cube_cells;
    std::vector<int64_t> cube_first;
    std::vector<size_t> whole;
    int64_t whole_first = 0;

    struct Piece {
        int64_t low, high;
        double open;
    };
    struct Square {
        size_t low = 0;
        int axis = 0;
        std::vector<Piece> pieces;
    };
    std::vector<Square> squares;

    struct Region {
        size_t cell = 0;
        double volume = 0;
        int64_t owner = kNone;
        bool in_air = false;
        std::vector<Wall> wall;
    };
    std::vector<Region> regions;
    std::vector<uint8_t> type;

    Fit *out = nullptr;

    // ---- indices -----------------------------------------------------------------------

    std::array<size_t, 3> at(size_t i) const { return {i % n[0], (i / n[0]) % n[1], i / sz}; }
    size_t index(size_t x, size_t y, size_t z) const { return x + y * sy + z * sz; }
    bool edge(size_t i) const {
        const auto p = at(i);
        for (int a = 0; a < 3; ++a)
            if (p[a] == 0 || p[a] + 1 == n[a]) return true;
        return false;
    }
    size_t depth(size_t i) const {
        const auto p = at(i);
        return std::min({p[0], p[1], p[2], n[0] - 1 - p[0], n[1] - 1 - p[1], n[2] - 1 - p[2]});
    }
    size_t neighbour(size_t i, int k) const {
        return static_cast<size_t>(static_cast<ptrdiff_t>(i) + kSteps[k][0] + kSteps[k][1] * static_cast<ptrdiff_t>(sy) +
                                   kSteps[k][2] * static_cast<ptrdiff_t>(sz));
    }
    // The direction (0..17) from cell a to cell b, or -1 where b is no neighbour of a.
    int direction(size_t a, size_t b) const {
        const auto pa = at(a), pb = at(b);
        int d[3];
        for (int c = 0; c < 3; ++c) {
            const long long v = static_cast<long long>(pb[c]) - static_cast<long long>(pa[c]);
            if (v < -1 || v > 1) return -1;
            d[c] = static_cast<int>(v);
        }
        for (int k = 0; k < 18; ++k)
            if (kSteps[k][0] == d[0] && kSteps[k][1] == d[1] && kSteps[k][2] == d[2]) return k;
        return -1;
    }
    size_t distance(size_t a, size_t b) const {
        const auto pa = at(a), pb = at(b);
        size_t m = 0;
        for (int c = 0; c < 3; ++c) m = std::max(m, pa[c] > pb[c] ? pa[c] - pb[c] : pb[c] - pa[c]);
        return m;
    }
    Frame corner(size_t i) const {
        const auto p = at(i);
        Frame f;
        for (int a = 0; a < 3; ++a) f.c[a] = G[a][p[a]];
        return f;
    }
    const Cube *cube_at(size_t cell) const {
        const auto it = std::lower_bound(cube_cells.begin(), cube_cells.end(), cell);
        if (it == cube_cells.end() || *it != cell) return nullptr;
        return &cubes[static_cast<size_t>(it - cube_cells.begin())];
    }
    bool has_cube(size_t cell) const { return std::binary_search(cube_cells.begin(), cube_cells.end(), cell); }

    // ---- the lattice -------------------------------------------------------------------

    int64_t quantize(double x, int a) const {
        if (!std::isfinite(x)) throw Refused("fitted walls: a coordinate that is not finite");
        i128 v = 0;
        if (x != 0.0) {
            int e = 0;
            const double m = std::frexp(x, &e);
            const auto M = static_cast<int64_t>(std::ldexp(m, 53));  // x = M 2^(e - 53)
            const int k = e - 53 + s;
            if (k >= 0) {
                if (k > 70) throw Refused("fitted walls: a coordinate outside the lattice's frame");
                v = static_cast<i128>(M) * (static_cast<i128>(1) << k);
            } else {
                const int r = -k;
                const u128 u = mag(M);
                u128 q = 0, rem = u;
                if (r < 120) {
                    q = u >> r;
                    rem = u - (q << r);
                    const u128 half = static_cast<u128>(1) << (r - 1);
                    if (rem > half || (rem == half && (q & 1))) ++q;
                }
                v = M < 0 ? -static_cast<i128>(q) : static_cast<i128>(q);
            }
        }
        const i128 t = v - X0s[a];
        const i128 lim = static_cast<i128>(1) << 61;
        if (t >= lim || t <= -lim) throw Refused("fitted walls: a coordinate outside the lattice's frame");
        return static_cast<int64_t>(t);
    }

    void frame(const Mesh &mesh, const GridSpec &g, std::vector<char> &given) {
        double glo[3], ghi[3], lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const double first = g.origin[a], last = g.origin[a] + static_cast<double>(n[a]) * h;
            glo[a] = first - 2.0 * h;
            ghi[a] = last + 2.0 * h;
            lo[a] = 1e300;
            hi[a] = -1e300;
        }
        given.assign(mesh.face_count, 0);
        for (size_t f = 0; f < mesh.face_count; ++f) {
            double p[3][3];
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) p[k][a] = mesh.vertices[3 * static_cast<size_t>(mesh.faces[3 * f + k]) + a];
            bool meets = true;
            for (int a = 0; a < 3 && meets; ++a) {
                const double mn = std::min({p[0][a], p[1][a], p[2][a]}), mx = std::max({p[0][a], p[1][a], p[2][a]});
                meets = mx >= glo[a] && mn <= ghi[a];
            }
            if (!meets) continue;
            given[f] = 1;
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], p[k][a]);
                    hi[a] = std::max(hi[a], p[k][a]);
                }
        }
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], g.origin[a]);
            hi[a] = std::max(hi[a], g.origin[a] + static_cast<double>(n[a]) * h);
        }
        double E = 0.0;
        for (int a = 0; a < 3; ++a) E = std::max(E, (hi[a] - lo[a]) / 2.0);
        E += 2.0 * h;
        if (E / h >= std::ldexp(1.0, 24)) {
            throw Refused("fitted walls: a face reaching the grid lies 2^24 cells or more from the middle of the "
                          "faces reaching it; split the face.");
        }
        int e = 0;
        const double m = std::frexp(E, &e);
        const int p = (m > 0.5 ? e : e - 1) + 1;
        s = 61 - p;
        for (int a = 0; a < 3; ++a) {
            const double X0 = std::ldexp(std::nearbyint(std::ldexp((lo[a] + hi[a]) / 2.0, -(p - 2))), p - 2);
            X0s[a] = static_cast<i128>(std::nearbyint(std::ldexp(X0, -(p - 2)))) * (static_cast<i128>(1) << 59);
        }
        S = std::ldexp(h, s);
        double reach = 0.0;
        for (int a = 0; a < 3; ++a)
            for (size_t i = 0; i <= n[a]; ++i) reach = std::max(reach, std::abs(g.origin[a] + static_cast<double>(i) * h));
        if (std::nextafter(reach, 1e308) - reach >= h / 4.0) {
            throw Refused("fitted walls: the grid lies too far from the origin of its coordinates for its cell");
        }
    }

    void lattice_grid(const GridSpec &g) {
        for (int a = 0; a < 3; ++a) {
            G[a].clear();
            C[a].clear();
            for (size_t i = 0; i <= n[a]; ++i) G[a].push_back(quantize(g.origin[a] + static_cast<double>(i) * h, a));
            for (size_t i = 0; i < n[a]; ++i)
                C[a].push_back(quantize(g.origin[a] + (static_cast<double>(i) + 0.5) * h, a));
            auto &w = out->width[a];
            w.clear();
            for (size_t i = 0; i < n[a]; ++i) {
                if (!(G[a][i] < C[a][i] && C[a][i] < G[a][i + 1]))
                    throw std::logic_error("fitted walls: a centre not inside its cell on the lattice");
                w.push_back(static_cast<double>(G[a][i + 1] - G[a][i]) / S);
                if (std::abs(w.back() - 1.0) > std::ldexp(1.0, -24)) {
                    throw Refused("fitted walls: the grid's cells along " + std::string(1, "xyz"[a]) +
                                  " are not whole cells on the lattice (a width off 1 by more than 2^-24): the grid "
                                  "lies too far from its origin for its cell");
                }
            }
        }
    }

    // ---- the mesh ----------------------------------------------------------------------

    void mesh_on_lattice(const Mesh &mesh, const int64_t *face_material, const std::vector<char> &given) {
        // Weld: one vertex per lattice point.
        std::vector<std::array<int64_t, 3>> q(mesh.vertex_count);
        std::vector<char> used(mesh.vertex_count, 0);
        for (size_t f = 0; f < mesh.face_count; ++f)
            if (given[f])
                for (int k = 0; k < 3; ++k) used[static_cast<size_t>(mesh.faces[3 * f + k])] = 1;
        std::vector<std::pair<std::array<int64_t, 3>, size_t>> keyed;
        for (size_t v = 0; v < mesh.vertex_count; ++v) {
            if (!used[v]) continue;
            for (int a = 0; a < 3; ++a) q[v][a] = quantize(mesh.vertices[3 * v + a], a);
            keyed.push_back({q[v], v});
        }
        std::sort(keyed.begin(), keyed.end());
        std::vector<int> wid(mesh.vertex_count, -1);
        for (size_t k = 0; k < keyed.size(); ++k) {
            if (k == 0 || keyed[k].first != keyed[k - 1].first) wq.push_back(keyed[k].first);
            wid[keyed[k].second] = static_cast<int>(wq.size() - 1);
        }
        // Faces and their planes.
        struct Key {
            i128 n[3];
            I256 d;
            size_t face;
        };
        std::vector<Key> keys;
        for (size_t f = 0; f < mesh.face_count; ++f) {
            if (!given[f]) continue;
            FaceX F;
            for (int k = 0; k < 3; ++k) F.v[k] = wid[static_cast<size_t>(mesh.faces[3 * f + k])];
            F.material = static_cast<int>(face_material[f]);
            F.index = static_cast<int64_t>(f);
            const auto &A = wq[F.v[0]], &B = wq[F.v[1]], &Cv = wq[F.v[2]];
            i128 e1[3], e2[3], nn[3];
            for (int a = 0; a < 3; ++a) {
                e1[a] = static_cast<i128>(B[a]) - A[a];
                e2[a] = static_cast<i128>(Cv[a]) - A[a];
            }
            nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
            nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
            nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
            if (nn[0] == 0 && nn[1] == 0 && nn[2] == 0) continue;  // no area: not a face of the fit
            const u128 g = gcd(gcd(mag(nn[0]), mag(nn[1])), mag(nn[2]));
            for (auto &c : nn) c /= static_cast<i128>(g);
            const int lead = nn[0] != 0 ? 0 : (nn[1] != 0 ? 1 : 2);
            if (nn[lead] < 0)
                for (auto &c : nn) c = -c;
            Key key;
            for (int a = 0; a < 3; ++a) key.n[a] = nn[a];
            key.d = mul(nn[0], A[0]) + mul(nn[1], A[1]) + mul(nn[2], A[2]);
            key.face = faces.size();
            for (int a = 0; a < 3; ++a) {
                F.lo[a] = std::min({A[a], B[a], Cv[a]});
                F.hi[a] = std::max({A[a], B[a], Cv[a]});
            }
            faces.push_back(F);
            keys.push_back(key);
        }
        if (faces.empty()) throw Refused("fitted walls: no face of the model reaches the grid");
        auto less = [](const Key &x, const Key &y) {
            for (int a = 0; a < 3; ++a)
                if (x.n[a] != y.n[a]) return x.n[a] < y.n[a];
            if (!(x.d == y.d)) return x.d < y.d;
            return x.face < y.face;
        };
        std::sort(keys.begin(), keys.end(), less);
        for (size_t k = 0; k < keys.size(); ++k) {
            const bool fresh = k == 0 || !(keys[k].n[0] == keys[k - 1].n[0] && keys[k].n[1] == keys[k - 1].n[1] &&
                                           keys[k].n[2] == keys[k - 1].n[2] && keys[k].d == keys[k - 1].d);
            if (fresh) {
                PlaneX P;
                for (int a = 0; a < 3; ++a) P.n[a] = keys[k].n[a];
                P.d = keys[k].d;
                const real len = std::sqrt(static_cast<real>(P.n[0]) * static_cast<real>(P.n[0]) +
                                           static_cast<real>(P.n[1]) * static_cast<real>(P.n[1]) +
                                           static_cast<real>(P.n[2]) * static_cast<real>(P.n[2]));
                for (int a = 0; a < 3; ++a) P.unit[a] = static_cast<double>(static_cast<real>(P.n[a]) / len);
                planes.push_back(P);
            }
            faces[keys[k].face].plane = static_cast<int>(planes.size() - 1);
        }
        merge_near_planes();
        // Membranes are not fitted here.
        for (const PlaneX &P : planes) {
            int axis = -1, count = 0;
            for (int a = 0; a < 3; ++a)
                if (P.n[a] != 0) {
                    axis = a;
                    ++count;
                }
            if (count != 1 || (P.n[axis] != 1 && P.n[axis] != -1)) continue;
            // d fits in 64 bits when its upper words are its sign's extension.
            const uint64_t fill = (P.d.w[0] >> 63) ? ~uint64_t{0} : 0;
            if (P.d.w[1] != fill || P.d.w[2] != fill || P.d.w[3] != fill) continue;
            const auto g = static_cast<int64_t>(P.d.w[0]) * (P.n[axis] > 0 ? 1 : -1);
            if (std::binary_search(G[axis].begin(), G[axis].end(), g)) {
                throw Refused("fitted walls: a face lies in a plane of the grid's cell faces; this set-up fits faces "
                              "off those planes only. Move the model or the grid by a fraction of a cell.");
            }
        }
        // Edges; the mesh must close inside the grid.
        std::vector<std::vector<int>> around(wq.size());
        for (size_t f = 0; f < faces.size(); ++f)
            for (int k = 0; k < 3; ++k) {
                const int a = faces[f].v[k], b = faces[f].v[(k + 1) % 3];
                edges[{std::min(a, b), std::max(a, b)}].push_back(static_cast<int>(f));
                around[static_cast<size_t>(faces[f].v[k])].push_back(static_cast<int>(f));
            }
        for (const auto &[e, fs] : edges) {
            if (fs.size() % 2 == 0) continue;
            bool inside = true;
            for (int a = 0; a < 3 && inside; ++a) {
                const int64_t lo = std::min(wq[e.first][a], wq[e.second][a]), hi = std::max(wq[e.first][a], wq[e.second][a]);
                inside = hi >= G[a].front() && lo <= G[a].back();
            }
            if (inside) {
                throw Refused("fitted walls: the model's faces do not close inside the grid (an edge met by an odd "
                              "count of faces); this set-up fits closed models only.");
            }
        }
        // The vertices on each plane: those of its faces, and any vertex of a face next to
        // them that lies on it exactly.
        std::vector<std::vector<int>> plane_faces(planes.size());
        for (size_t f = 0; f < faces.size(); ++f) plane_faces[static_cast<size_t>(faces[f].plane)].push_back(static_cast<int>(f));
        for (size_t p = 0; p < planes.size(); ++p) {
            std::vector<int> cand;
            for (const int f : plane_faces[p])
                for (const int v : faces[static_cast<size_t>(f)].v)
                    for (const int g2 : around[static_cast<size_t>(v)])
                        for (const int w : faces[static_cast<size_t>(g2)].v) cand.push_back(w);
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            std::vector<int> own;
            for (const int f : plane_faces[p])
                for (const int v : faces[static_cast<size_t>(f)].v) own.push_back(v);
            std::sort(own.begin(), own.end());
            for (const int w : cand)
                if (std::binary_search(own.begin(), own.end(), w) ||
                    on_plane_exact(static_cast<int>(p), wq[static_cast<size_t>(w)].data()) == 0)
                    planes[p].verts.push_back(w);
        }
        // Vertices on the grid's planes.
        for (size_t w = 0; w < wq.size(); ++w)
            for (int a = 0; a < 3; ++a) {
                const auto it = std::lower_bound(G[a].begin(), G[a].end(), wq[w][a]);
                if (it != G[a].end() && *it == wq[w][a]) axis_verts[{a, static_cast<size_t>(it - G[a].begin())}].push_back(static_cast<int>(w));
            }
        // Coverage normals.
        for (FaceX &F : faces) {
            const PlaneX &P = planes[static_cast<size_t>(F.plane)];
            for (int k = 0; k < 3; ++k) {
                const auto &A = wq[static_cast<size_t>(F.v[k])], &B = wq[static_cast<size_t>(F.v[(k + 1) % 3])];
                const auto &Cv = wq[static_cast<size_t>(F.v[(k + 2) % 3])];
                real e[3], c[3], t[3];
                for (int a = 0; a < 3; ++a) {
                    e[a] = static_cast<real>(B[a] - A[a]);
                    t[a] = static_cast<real>(Cv[a] - A[a]);
                }
                c[0] = e[1] * P.unit[2] - e[2] * P.unit[1];
                c[1] = e[2] * P.unit[0] - e[0] * P.unit[2];
                c[2] = e[0] * P.unit[1] - e[1] * P.unit[0];
                const real len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                const real towards = c[0] * t[0] + c[1] * t[1] + c[2] * t[2];
                for (int a = 0; a < 3; ++a) F.cover_normal[k][a] = static_cast<double>((towards < 0 ? -c[a] : c[a]) / len);
            }
        }
        out->stats.faces = faces.size();
        out->stats.planes = planes.size();
        out->stats.quantum_bits = s;
    }

    // The distance of lattice point q from face plane p, in cells (exact, then rounded).
    real distance_to(int p, const int64_t *q) const {
        const PlaneX &P = planes[static_cast<size_t>(p)];
        const real len = std::sqrt(static_cast<real>(P.n[0]) * static_cast<real>(P.n[0]) + static_cast<real>(P.n[1]) * static_cast<real>(P.n[1]) +
                                   static_cast<real>(P.n[2]) * static_cast<real>(P.n[2]));
        return to_real(mul(P.n[0], q[0]) + mul(P.n[1], q[1]) + mul(P.n[2], q[2]) - P.d) / (len * static_cast<real>(S));
    }

    // Faces of one flat part of the model whose lattice points no longer lie on one plane
    // (the rounding onto the lattice tilts them by about 1e-16 of a cell) are taken on one
    // plane, the first of theirs: planes joined by a welded edge, every point of each face
    // within 1e-9 of a cell of the other's plane. Their faces then meet exactly along their
    // edges here; the reference decides the tiny wedges between such planes exactly, which
    // moves no decision and changes its measures by about 1e-16.
    void merge_near_planes() {
        const size_t P = planes.size();
        UF uf(P);
        auto near_all = [&](int p, const FaceX &F) {
            for (const int w : F.v)
                if (std::abs(distance_to(p, wq[static_cast<size_t>(w)].data())) > 1e-9) return false;
            return true;
        };
        std::map<std::pair<int, int>, std::vector<int>> shared;
        for (size_t f = 0; f < faces.size(); ++f)
            for (int k = 0; k < 3; ++k) {
                const int a = faces[f].v[k], b = faces[f].v[(k + 1) % 3];
                shared[{std::min(a, b), std::max(a, b)}].push_back(static_cast<int>(f));
            }
        for (const auto &[e, fs] : shared)
            for (size_t i = 0; i < fs.size(); ++i)
                for (size_t j = i + 1; j < fs.size(); ++j) {
                    const FaceX &F = faces[static_cast<size_t>(fs[i])], &H = faces[static_cast<size_t>(fs[j])];
                    if (F.plane == H.plane || uf.find(static_cast<size_t>(F.plane)) == uf.find(static_cast<size_t>(H.plane))) continue;
                    if (near_all(F.plane, H) && near_all(H.plane, F)) uf.join(static_cast<size_t>(F.plane), static_cast<size_t>(H.plane));
                }
        std::vector<int> keep(P, -1);
        std::vector<PlaneX> kept;
        for (size_t p = 0; p < P; ++p)
            if (uf.find(p) == p) {
                keep[p] = static_cast<int>(kept.size());
                kept.push_back(planes[p]);
            }
        if (kept.size() == P) return;
        for (FaceX &F : faces) {
            const int rep = static_cast<int>(uf.find(static_cast<size_t>(F.plane)));
            if (!near_all(rep, F)) {
                throw Refused("fitted walls: a flat part of the model whose faces the lattice tilts apart by more than "
                              "1e-9 of a cell; split it into faces on one plane each.");
            }
            F.plane = keep[static_cast<size_t>(rep)];
        }
        planes = std::move(kept);
    }

    // The sign of lattice point q against face plane p (exact).
    int on_plane_exact(int p, const int64_t *q) const {
        const PlaneX &P = planes[static_cast<size_t>(p)];
        return sign(mul(P.n[0], q[0]) + mul(P.n[1], q[1]) + mul(P.n[2], q[2]) - P.d);
    }

    // ---- incidence ---------------------------------------------------------------------

    bool has_vertex(int64_t id, int w) const {
        if (id < kAxisPlane) return std::binary_search(planes[static_cast<size_t>(id)].verts.begin(), planes[static_cast<size_t>(id)].verts.end(), w);
        if (id < kCoverPlane) {
            const int64_t v = id - kAxisPlane;
            const int a = static_cast<int>(v % 3);
            const auto i = static_cast<size_t>(v / 3);
            return wq[static_cast<size_t>(w)][a] == G[a][i];
        }
        const int64_t v = id - kCoverPlane;
        const FaceX &F = faces[static_cast<size_t>(v / 3)];
        const int k = static_cast<int>(v % 3);
        return F.v[k] == w || F.v[(k + 1) % 3] == w;
    }

    // Mesh vertices a plane holds by construction (a coverage plane: its edge's ends).
    void vertices_of(int64_t id, std::vector<int> &outv) const {
        outv.clear();
        if (id < kAxisPlane) {
            outv = planes[static_cast<size_t>(id)].verts;
        } else if (id < kCoverPlane) {
            const int64_t v = id - kAxisPlane;
            const auto it = axis_verts.find({static_cast<int>(v % 3), static_cast<size_t>(v / 3)});
            if (it != axis_verts.end()) outv = it->second;
        } else {
            const int64_t v = id - kCoverPlane;
            const FaceX &F = faces[static_cast<size_t>(v / 3)];
            const int k = static_cast<int>(v % 3);
            outv = {F.v[k], F.v[(k + 1) % 3]};
        }
    }

    void normal_of(int64_t id, double nrm[3]) const {
        if (id < kAxisPlane) {
            for (int a = 0; a < 3; ++a) nrm[a] = planes[static_cast<size_t>(id)].unit[a];
        } else if (id < kCoverPlane) {
            const int a = static_cast<int>((id - kAxisPlane) % 3);
            for (int c = 0; c < 3; ++c) nrm[c] = c == a ? 1.0 : 0.0;
        } else {
            const int64_t v = id - kCoverPlane;
            for (int a = 0; a < 3; ++a) nrm[a] = faces[static_cast<size_t>(v / 3)].cover_normal[v % 3][a];
        }
    }

    // Whether planes a and b (not parallel) meet in a line through two mesh vertices that
    // plane t holds too.
    bool line_in(int64_t a, int64_t b, int64_t t) const {
        double na[3], nb[3];
        normal_of(a, na);
        normal_of(b, nb);
        const double c0 = na[1] * nb[2] - na[2] * nb[1], c1 = na[2] * nb[0] - na[0] * nb[2], c2 = na[0] * nb[1] - na[1] * nb[0];
        if (c0 * c0 + c1 * c1 + c2 * c2 < 1e-24) return false;
        thread_local std::vector<int> va;
        vertices_of(a, va);
        int found = 0;
        for (const int w : va) {
            if (has_vertex(b, w) && has_vertex(t, w) && ++found == 2) return true;
        }
        return false;
    }

    bool incident(const Vx &v, int64_t id) const {
        if (v.has(id)) return true;
        if (v.vid >= 0 && has_vertex(id, v.vid)) return true;
        for (int i = 0; i < v.n; ++i)
            for (int j = i + 1; j < v.n; ++j)
                if (line_in(v.on[i], v.on[j], id)) return true;
        return false;
    }

    std::string plane_text(int64_t id) const {
        if (id < kAxisPlane) return "face plane " + std::to_string(id);
        if (id < kCoverPlane) return "grid plane " + std::string(1, "xyz"[(id - kAxisPlane) % 3]) + " " + std::to_string((id - kAxisPlane) / 3);
        return "edge " + std::to_string((id - kCoverPlane) % 3) + " of face " + std::to_string(faces[static_cast<size_t>((id - kCoverPlane) / 3)].index);
    }

    static real dist(const Vx &v, const LP &p) { return p.un[0] * v.u[0] + p.un[1] * v.u[1] + p.un[2] * v.u[2] - p.off; }

    int side(const Vx &v, const LP &p, real &d) const {
        if (incident(v, p.id)) {
            d = 0;
            return 0;
        }
        d = dist(v, p);
        if (std::abs(d) <= kTol) {
            std::string what = "a point of the model's cut (vertex " + std::to_string(v.vid) + ", planes";
            for (int k = 0; k < v.n; ++k) what += " " + plane_text(v.on[k]);
            what += ") meets plane " + plane_text(p.id) + " at " + std::to_string(static_cast<double>(d));
            if (std::getenv("AS_FIT_DEBUG")) {
                std::fprintf(stderr, "point u %.17Lg %.17Lg %.17Lg\n", v.u[0], v.u[1], v.u[2]);
                for (int k = 0; k < v.n; ++k) {
                    double nn[3];
                    normal_of(v.on[k], nn);
                    std::fprintf(stderr, "  on %s normal %.6f %.6f %.6f\n", plane_text(v.on[k]).c_str(), nn[0], nn[1], nn[2]);
                }
                double nn[3];
                normal_of(p.id, nn);
                std::fprintf(stderr, "  against %s normal %.6f %.6f %.6f off %.17Lg\n", plane_text(p.id).c_str(), nn[0], nn[1], nn[2], p.off);
                for (int k = 0; k < v.n; ++k) {
                    std::vector<int> vv;
                    vertices_of(v.on[k], vv);
                    std::fprintf(stderr, "  verts of %s:", plane_text(v.on[k]).c_str());
                    for (int w : vv) std::fprintf(stderr, " %d", w);
                    std::fprintf(stderr, "\n");
                }
                std::vector<int> vv;
                vertices_of(p.id, vv);
                std::fprintf(stderr, "  verts of against:");
                for (int w : vv) std::fprintf(stderr, " %d", w);
                std::fprintf(stderr, "\n");
            }
            too_close(what.c_str());
        }
        return d > 0 ? 1 : -1;
    }

    // ---- planes and points in a frame ---------------------------------------------------

    // A lattice coordinate in cells from the frame's corner: the double nearest the
    // difference, over S, rounded to a double (as each cell's width is formed), then
    // carried in long double.
    real local(const Frame &fr, int a, int64_t q) const { return static_cast<real>(static_cast<double>(q - fr.c[a]) / S); }
    // A mesh vertex's coordinate in the frame, in long double (the difference exact).
    real local_exact(const Frame &fr, int a, int64_t q) const { return static_cast<real>(q - fr.c[a]) / static_cast<real>(S); }

    LP face_lp(int p, const Frame &fr) const {
        const PlaneX &P = planes[static_cast<size_t>(p)];
        LP l;
        l.id = p;
        const real len = std::sqrt(static_cast<real>(P.n[0]) * static_cast<real>(P.n[0]) + static_cast<real>(P.n[1]) * static_cast<real>(P.n[1]) +
                                   static_cast<real>(P.n[2]) * static_cast<real>(P.n[2]));
        for (int a = 0; a < 3; ++a) l.un[a] = static_cast<real>(P.n[a]) / len;
        const I256 rel = P.d - (mul(P.n[0], fr.c[0]) + mul(P.n[1], fr.c[1]) + mul(P.n[2], fr.c[2]));
        l.off = to_real(rel) / (len * static_cast<real>(S));
        return l;
    }

    LP axis_lp(int a, size_t i, const Frame &fr) const {
        LP l;
        l.id = axis_id(a, i);
        l.un[a] = 1;
        l.off = local(fr, a, G[a][i]);
        return l;
    }

    LP cover_lp(size_t f, int k, const Frame &fr) const {
        const FaceX &F = faces[f];
        const PlaneX &P = planes[static_cast<size_t>(F.plane)];
        const auto &A = wq[static_cast<size_t>(F.v[k])], &B = wq[static_cast<size_t>(F.v[(k + 1) % 3])];
        const auto &Cv = wq[static_cast<size_t>(F.v[(k + 2) % 3])];
        const real len = std::sqrt(static_cast<real>(P.n[0]) * static_cast<real>(P.n[0]) + static_cast<real>(P.n[1]) * static_cast<real>(P.n[1]) +
                                   static_cast<real>(P.n[2]) * static_cast<real>(P.n[2]));
        real nu[3], e[3], t[3], c[3];
        for (int a = 0; a < 3; ++a) {
            nu[a] = static_cast<real>(P.n[a]) / len;
            e[a] = static_cast<real>(B[a] - A[a]);
            t[a] = static_cast<real>(Cv[a] - A[a]);
        }
        c[0] = e[1] * nu[2] - e[2] * nu[1];
        c[1] = e[2] * nu[0] - e[0] * nu[2];
        c[2] = e[0] * nu[1] - e[1] * nu[0];
        const real cl = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const real sgn = (c[0] * t[0] + c[1] * t[1] + c[2] * t[2]) < 0 ? -1 : 1;
        LP l;
        l.id = cover_id(f, k);
        real off = 0;
        for (int a = 0; a < 3; ++a) {
            l.un[a] = sgn * c[a] / cl;
            off += l.un[a] * local_exact(fr, a, A[a]);
        }
        l.off = off;
        return l;
    }

    Vx vertex(int w, const Frame &fr) const {
        Vx v;
        v.vid = w;
        for (int a = 0; a < 3; ++a) v.u[a] = local_exact(fr, a, wq[static_cast<size_t>(w)][a]);
        return v;
    }

    // A corner of the grid's cells (lattice indices i per axis) in a frame.
    Vx grid_corner(const size_t idx[3], const Frame &fr) const {
        Vx v;
        for (int a = 0; a < 3; ++a) {
            v.u[a] = local(fr, a, G[a][idx[a]]);
            v.add(axis_id(a, idx[a]));
        }
        return v;
    }

    // The point where the segment a b meets plane p (da, db: their distances, of either sign).
    Vx meet(const Vx &a, const Vx &b, real da, real db, const LP &p, const Frame &fr) const {
        Vx v;
        const real t = da / (da - db);
        for (int c = 0; c < 3; ++c) v.u[c] = a.u[c] + t * (b.u[c] - a.u[c]);
        v.add(p.id);
        for (int k = 0; k < a.n; ++k)
            if (incident(b, a.on[k])) v.add(a.on[k]);
        for (int k = 0; k < b.n; ++k)
            if (incident(a, b.on[k])) v.add(b.on[k]);
        if (a.vid >= 0 && b.vid >= 0) {
            const auto it = edges.find({std::min(a.vid, b.vid), std::max(a.vid, b.vid)});
            if (it != edges.end())
                for (const int f : it->second) {
                    const FaceX &F = faces[static_cast<size_t>(f)];
                    v.add(F.plane);
                    for (int k = 0; k < 3; ++k) {
                        const int x = F.v[k], y = F.v[(k + 1) % 3];
                        if ((x == a.vid && y == b.vid) || (x == b.vid && y == a.vid)) v.add(cover_id(static_cast<size_t>(f), k));
                    }
                }
        }
        snap(v, fr);
        return v;
    }

    // A point made on three planes that meet in one point only, a mesh vertex on all
    // three, is that vertex.
    void snap(Vx &v, const Frame &fr) const {
        if (v.n < 3 || v.vid >= 0) return;
        thread_local std::vector<int> held;
        for (int i = 0; i < v.n; ++i)
            for (int j = i + 1; j < v.n; ++j)
                for (int k = j + 1; k < v.n; ++k) {
                    double a[3], b[3], c[3];
                    normal_of(v.on[i], a);
                    normal_of(v.on[j], b);
                    normal_of(v.on[k], c);
                    const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                                       a[2] * (b[0] * c[1] - b[1] * c[0]);
                    if (std::abs(det) < 1e-9) continue;
                    vertices_of(v.on[i], held);
                    for (const int w : held) {
                        if (has_vertex(v.on[j], w) && has_vertex(v.on[k], w)) {
                            const Vx x = vertex(w, fr);
                            v.vid = w;
                            for (int c2 = 0; c2 < 3; ++c2) v.u[c2] = x.u[c2];
                            return;
                        }
                    }
                }
    }

    // Keeps the part of g on side `keep` of p (closed); false, g emptied, where no part of
    // positive area is left (no point strictly on the kept side).
    bool clip(Poly &g, const LP &p, int keep, const Frame &fr) const {
        thread_local std::vector<int> sd;
        thread_local std::vector<real> ds;
        const size_t m = g.size();
        sd.resize(m);
        ds.resize(m);
        bool in = false, outside = false;
        for (size_t i = 0; i < m; ++i) {
            sd[i] = side(g[i], p, ds[i]) * keep;
            in = in || sd[i] > 0;
            outside = outside || sd[i] < 0;
        }
        if (!in) {
            g.clear();
            return false;
        }
        if (!outside) return true;
        Poly r;
        r.reserve(m + 2);
        for (size_t i = 0; i < m; ++i) {
            const size_t j = (i + 1) % m;
            if (sd[i] >= 0) r.push_back(g[i]);
            if ((sd[i] > 0 && sd[j] < 0) || (sd[i] < 0 && sd[j] > 0)) r.push_back(meet(g[i], g[j], ds[i], ds[j], p, fr));
        }
        g = std::move(r);
        return true;
    }

    // The area of planar polygon g whose plane has unit normal un.
    static real area(const Poly &g, const real un[3]) {
        if (g.size() < 3) return 0;
        int a = 0;
        for (int c = 1; c < 3; ++c)
            if (std::abs(un[c]) > std::abs(un[a])) a = c;
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        real sum = 0;
        const Vx &o = g[0];
        for (size_t i = 1; i + 1 < g.size(); ++i) {
            const real x1 = g[i].u[b] - o.u[b], y1 = g[i].u[c] - o.u[c];
            const real x2 = g[i + 1].u[b] - o.u[b], y2 = g[i + 1].u[c] - o.u[c];
            sum += x1 * y2 - x2 * y1;
        }
        return std::abs(sum) / (2 * std::abs(un[a]));
    }

    // The polygon of plane p inside box [0, w] of frame fr (cell `cell`'s cube).
    Poly plane_in_box(const LP &p, size_t cell, const Frame &fr, const LP box[6]) const {
        const auto idx = at(cell);
        real cen[3];
        for (int a = 0; a < 3; ++a) cen[a] = local(fr, a, G[a][idx[a] + 1]) / 2;
        const real d = p.un[0] * cen[0] + p.un[1] * cen[1] + p.un[2] * cen[2] - p.off;
        real o[3], e1[3], e2[3];
        for (int a = 0; a < 3; ++a) o[a] = cen[a] - d * p.un[a];
        int m = 0;
        for (int a = 1; a < 3; ++a)
            if (std::abs(p.un[a]) < std::abs(p.un[m])) m = a;
        real ax[3] = {0, 0, 0};
        ax[m] = 1;
        e1[0] = p.un[1] * ax[2] - p.un[2] * ax[1];
        e1[1] = p.un[2] * ax[0] - p.un[0] * ax[2];
        e1[2] = p.un[0] * ax[1] - p.un[1] * ax[0];
        const real l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (auto &c : e1) c /= l1;
        e2[0] = p.un[1] * e1[2] - p.un[2] * e1[1];
        e2[1] = p.un[2] * e1[0] - p.un[0] * e1[2];
        e2[2] = p.un[0] * e1[1] - p.un[1] * e1[0];
        Poly g(4);
        const real sx[4] = {-3, 3, 3, -3}, sy2[4] = {-3, -3, 3, 3};
        for (int k = 0; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) g[static_cast<size_t>(k)].u[a] = o[a] + sx[k] * e1[a] + sy2[k] * e2[a];
            g[static_cast<size_t>(k)].add(p.id);
        }
        for (int k = 0; k < 6; ++k)
            if (!clip(g, box[k], (k & 1) ? -1 : 1, fr)) return {};
        return g;
    }

    // The box of cell i in frame fr: for each axis its low plane (keep +) and high plane
    // (keep -), in that order.
    void box_of(size_t cell, const Frame &fr, LP box[6]) const {
        const auto idx = at(cell);
        for (int a = 0; a < 3; ++a) {
            box[2 * a] = axis_lp(a, idx[a], fr);
            box[2 * a + 1] = axis_lp(a, idx[a] + 1, fr);
        }
    }

    // The square of cell i's side k (2 a: low, 2 a + 1: high along a) in frame fr.
    Poly square_of(size_t cell, int k, const Frame &fr) const {
        const auto idx = at(cell);
        const int a = k / 2, b = (a + 1) % 3, c = (a + 2) % 3;
        Poly g;
        const int corners[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (const auto &cr : corners) {
            size_t ix[3];
            ix[a] = idx[a] + static_cast<size_t>(k & 1);
            ix[b] = idx[b] + static_cast<size_t>(cr[0]);
            ix[c] = idx[c] + static_cast<size_t>(cr[1]);
            g.push_back(grid_corner(ix, fr));
        }
        return g;
    }

    Poly triangle(size_t f, const Frame &fr) const {
        Poly g;
        for (int k = 0; k < 3; ++k) {
            Vx v = vertex(faces[f].v[k], fr);
            v.add(faces[f].plane);
            g.push_back(v);
        }
        return g;
    }

    // ---- faces into cubes ---------------------------------------------------------------

    // Cells along axis a whose closed extent meets [lo, hi], within 1 .. n - 2.
    bool cell_range(int a, int64_t lo, int64_t hi, size_t &i0, size_t &i1) const {
        const auto &g = G[a];
        // first cell i with G[i + 1] >= lo
        size_t first = static_cast<size_t>(std::lower_bound(g.begin(), g.end(), lo) - g.begin());
        first = first == 0 ? 0 : first - 1;
        size_t last = static_cast<size_t>(std::upper_bound(g.begin(), g.end(), hi) - g.begin());
        if (last == 0) return false;
        last -= 1;  // G[last] <= hi
        i0 = std::max<size_t>(first, 1);
        i1 = std::min<size_t>(last, n[a] - 2);
        if (last >= n[a]) i1 = n[a] - 2;
        return i0 <= i1;
    }

    void faces_into_cubes() {
        std::vector<std::vector<std::pair<size_t, int>>> found(static_cast<size_t>(threads));
        const size_t T = static_cast<size_t>(std::max(threads, 1));
        found.resize(T);
        std::vector<std::exception_ptr> errors(T);
        auto run = [&](size_t t) {
            try {
                for (size_t f = t; f < faces.size(); f += T) {
                    const FaceX &F = faces[f];
                    size_t r0[3], r1[3];
                    bool any = true;
                    for (int a = 0; a < 3 && any; ++a) any = cell_range(a, F.lo[a], F.hi[a], r0[a], r1[a]);
                    if (!any) continue;
                    const PlaneX &P = planes[static_cast<size_t>(F.plane)];
                    for (size_t z = r0[2]; z <= r1[2]; ++z)
                        for (size_t y = r0[1]; y <= r1[1]; ++y)
                            for (size_t x = r0[0]; x <= r1[0]; ++x) {
                                const size_t cell = index(x, y, z);
                                // A quick test: the plane's distance from the cube's centre.
                                double cen = 0, reach = 0;
                                const size_t ix[3] = {x, y, z};
                                for (int a = 0; a < 3; ++a) {
                                    const double mid = (static_cast<double>(G[a][ix[a]]) + static_cast<double>(G[a][ix[a] + 1])) / 2.0;
                                    cen += P.unit[a] * (mid - static_cast<double>(wq[static_cast<size_t>(F.v[0])][a]));
                                    reach += std::abs(P.unit[a]) * (static_cast<double>(G[a][ix[a] + 1]) - static_cast<double>(G[a][ix[a]])) / 2.0;
                                }
                                if (std::abs(cen) > reach * (1.0 + 1e-6) + 1e-3 * S) continue;
                                const Frame fr = corner(cell);
                                LP box[6];
                                box_of(cell, fr, box);
                                Poly g = triangle(f, fr);
                                bool meets = true;
                                for (int k = 0; k < 6 && meets; ++k) meets = clip(g, box[k], (k & 1) ? -1 : 1, fr);
                                if (meets) found[t].push_back({cell, static_cast<int>(f)});
                            }
                }
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
        std::vector<std::pair<size_t, int>> all;
        for (auto &v : found) all.insert(all.end(), v.begin(), v.end());
        std::sort(all.begin(), all.end());
        for (size_t k = 0; k < all.size();) {
            Cube c;
            c.cell = all[k].first;
            for (; k < all.size() && all[k].first == c.cell; ++k) c.faces.push_back(all[k].second);
            for (const int f : c.faces) c.planes.push_back(faces[static_cast<size_t>(f)].plane);
            std::sort(c.planes.begin(), c.planes.end());
            c.planes.erase(std::unique(c.planes.begin(), c.planes.end()), c.planes.end());
            cube_cells.push_back(c.cell);
            cubes.push_back(std::move(c));
        }
        out->stats.cubes = cubes.size();
    }

    // ---- a cube ---------------------------------------------------------------------------

    // Which leaf of the cube holds lattice point q (a point on a plane on its + side).
    int leaf_holding(const Cube &c, const int64_t q[3]) const {
        for (size_t l = 0; l < c.leaves.size(); ++l) {
            bool inside = true;
            for (const Cut &cut : c.leaves[l]) {
                const int sd = on_plane_exact(cut.plane, q);
                inside = cut.side > 0 ? sd >= 0 : sd < 0;
                if (!inside) break;
            }
            if (inside) return static_cast<int>(l);
        }
        throw std::logic_error("fitted walls: a point in no leaf of its cube");
    }

    void cut_cube(Cube &c) const {
        const Frame fr = corner(c.cell);
        LP box[6];
        box_of(c.cell, fr, box);
        const size_t P = c.planes.size();
        std::vector<LP> lp(P);
        for (size_t k = 0; k < P; ++k) lp[k] = face_lp(c.planes[k], fr);
        auto local_of = [&](int plane) {
            return static_cast<size_t>(std::lower_bound(c.planes.begin(), c.planes.end(), plane) - c.planes.begin());
        };
        // Each face's part in the cube.
        std::vector<Poly> part(c.faces.size());
        for (size_t k = 0; k < c.faces.size(); ++k) {
            part[k] = triangle(static_cast<size_t>(c.faces[k]), fr);
            for (int b = 0; b < 6; ++b)
                if (!clip(part[k], box[b], (b & 1) ? -1 : 1, fr)) throw std::logic_error("fitted walls: a face left its cube");
        }
        // Leaves: split by each plane where a face on it has positive area inside.
        std::vector<Leaf> leaves{Leaf{}};
        for (size_t k = 0; k < P; ++k) {
            std::vector<Leaf> next;
            for (const Leaf &leaf : leaves) {
                bool hit = false;
                for (size_t f = 0; f < c.faces.size() && !hit; ++f) {
                    if (faces[static_cast<size_t>(c.faces[f])].plane != c.planes[k]) continue;
                    Poly g = part[f];
                    bool positive = true;
                    for (const Cut &cut : leaf) {
                        positive = clip(g, lp[local_of(cut.plane)], cut.side, fr);
                        if (!positive) break;
                    }
                    hit = positive;
                }
                if (!hit) {
                    next.push_back(leaf);
                    continue;
                }
                Leaf lo = leaf, hi = leaf;
                lo.push_back({c.planes[k], -1});
                hi.push_back({c.planes[k], 1});
                next.push_back(std::move(lo));
                next.push_back(std::move(hi));
            }
            leaves = std::move(next);
        }
        const size_t L = leaves.size();
        // Each plane's polygon in the box, and each leaf's facets on its planes.
        std::vector<Poly> in_box(P);
        for (size_t k = 0; k < P; ++k) in_box[k] = plane_in_box(lp[k], c.cell, fr, box);
        auto facet = [&](const Leaf &leaf, size_t skip) {
            const size_t k = local_of(leaf[skip].plane);
            Poly g = in_box[k];
            for (size_t j = 0; j < leaf.size() && !g.empty(); ++j) {
                if (j == skip) continue;
                clip(g, lp[local_of(leaf[j].plane)], leaf[j].side, fr);
            }
            return g;
        };
        // Volumes by the divergence theorem about the cube's centre.
        real cen[3];
        for (int a = 0; a < 3; ++a) cen[a] = box[2 * a + 1].off / 2;
        std::vector<double> leaf_volume(L, 0.0);
        for (size_t l = 0; l < L; ++l) {
            real v = 0;
            for (int k = 0; k < 6; ++k) {
                Poly g = square_of(c.cell, k, fr);
                bool positive = true;
                for (const Cut &cut : leaves[l]) {
                    positive = clip(g, lp[local_of(cut.plane)], cut.side, fr);
                    if (!positive) break;
                }
                if (!positive) continue;
                v += area(g, box[k].un) * cen[k / 2];
            }
            for (size_t j = 0; j < leaves[l].size(); ++j) {
                const Poly g = facet(leaves[l], j);
                if (g.empty()) continue;
                const LP &p = lp[local_of(leaves[l][j].plane)];
                const real h0 = p.un[0] * cen[0] + p.un[1] * cen[1] + p.un[2] * cen[2] - p.off;
                v += area(g, p.un) * (static_cast<real>(leaves[l][j].side) * h0);
            }
            leaf_volume[l] = static_cast<double>(v / 3);
        }
        // Regions: leaves on either side of a plane join through what its faces leave bare.
        UF uf(L);
        std::vector<std::vector<Wall>> leaf_wall(L);
        for (size_t k = 0; k < P; ++k) {
            const int plane = c.planes[k];
            for (size_t a = 0; a < L; ++a) {
                size_t ja = 0;
                while (ja < leaves[a].size() && !(leaves[a][ja].plane == plane && leaves[a][ja].side > 0)) ++ja;
                if (ja == leaves[a].size()) continue;
                const Poly fa = facet(leaves[a], ja);
                if (fa.empty()) continue;
                for (size_t b = 0; b < L; ++b) {
                    size_t jb = 0;
                    while (jb < leaves[b].size() && !(leaves[b][jb].plane == plane && leaves[b][jb].side < 0)) ++jb;
                    if (jb == leaves[b].size()) continue;
                    Poly o = fa;
                    bool positive = true;
                    for (size_t j = 0; j < leaves[b].size() && positive; ++j) {
                        if (j == jb) continue;
                        positive = clip(o, lp[local_of(leaves[b][j].plane)], leaves[b][j].side, fr);
                    }
                    if (!positive) continue;
                    // Cover the overlap by the plane's faces, in order.
                    std::vector<Poly> bare{o};
                    for (size_t f = 0; f < c.faces.size() && !bare.empty(); ++f) {
                        const size_t face = static_cast<size_t>(c.faces[f]);
                        if (faces[face].plane != plane) continue;
                        LP cov[3];
                        for (int e = 0; e < 3; ++e) cov[e] = cover_lp(face, e, fr);
                        std::vector<Poly> left;
                        for (Poly &piece : bare) {
                            Poly inside = std::move(piece);
                            bool gone = false;
                            for (int e = 0; e < 3; ++e) {
                                Poly outer = inside;
                                if (clip(outer, cov[e], -1, fr)) left.push_back(std::move(outer));
                                if (!clip(inside, cov[e], 1, fr)) {
                                    gone = true;
                                    break;
                                }
                            }
                            if (gone) continue;
                            const double covered = static_cast<double>(area(inside, lp[k].un));
                            leaf_wall[a].push_back({faces[face].material, covered});
                            leaf_wall[b].push_back({faces[face].material, covered});
                        }
                        bare = std::move(left);
                    }
                    if (!bare.empty()) uf.join(a, b);
                }
            }
        }
        // The regions, numbered by their first leaf.
        std::vector<int> root_region(L, -1);
        c.leaf_region.assign(L, -1);
        c.volume.clear();
        c.wall.clear();
        for (size_t l = 0; l < L; ++l) {
            const size_t r = uf.find(l);
            if (root_region[r] < 0) {
                root_region[r] = static_cast<int>(c.volume.size());
                c.volume.push_back(0.0);
                c.wall.emplace_back();
            }
            const int g = root_region[r];
            c.leaf_region[l] = g;
            c.volume[static_cast<size_t>(g)] += leaf_volume[l];
            for (const Wall &w : leaf_wall[l]) {
                auto &ws = c.wall[static_cast<size_t>(g)];
                auto it = std::find_if(ws.begin(), ws.end(), [&](const Wall &x) { return x.material == w.material; });
                if (it == ws.end()) ws.push_back(w);
                else it->area += w.area;
            }
        }
        c.leaves = std::move(leaves);
        c.leaf_count = L;
        int64_t q[3];
        const auto idx = at(c.cell);
        for (int a = 0; a < 3; ++a) q[a] = C[a][idx[a]];
        c.centre = c.leaf_region[static_cast<size_t>(leaf_holding(c, q))];
    }

    void cut_cubes() {
        parallel(cubes.size(), threads, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; ++k) cut_cube(cubes[k]);
        });
        size_t leaves = 0;
        for (const Cube &c : cubes) leaves += c.leaf_count;
        out->stats.leaves = leaves;
    }

    // ---- squares ----------------------------------------------------------------------------

    int64_t region_id(size_t cell, const Cube *c, size_t leaf) const {
        if (c) return cube_first[static_cast<size_t>(c - cubes.data())] + c->leaf_region[leaf];
        if (edge(cell)) return kEdge;
        const auto it = std::lower_bound(whole.begin(), whole.end(), cell);
        return whole_first + static_cast<int64_t>(it - whole.begin());
    }

    Square cut_square(size_t low, int axis) const {
        Square sq;
        sq.low = low;
        sq.axis = axis;
        const size_t high = low + (axis == 0 ? 1 : (axis == 1 ? sy : sz));
        const Cube *cl = cube_at(low), *ch = cube_at(high);
        const Frame fr = corner(low);
        auto lps = [&](const Cube *c) {
            std::vector<LP> v;
            if (c)
                for (const int p : c->planes) v.push_back(face_lp(p, fr));
            return v;
        };
        const std::vector<LP> ll = lps(cl), lh = lps(ch);
        auto lp_of = [&](const Cube *c, const std::vector<LP> &v, int plane) -> const LP & {
            return v[static_cast<size_t>(std::lower_bound(c->planes.begin(), c->planes.end(), plane) - c->planes.begin())];
        };
        const Poly square = square_of(low, 2 * axis + 1, fr);
        auto facets = [&](const Cube *c, const std::vector<LP> &v) {
            std::vector<std::pair<size_t, Poly>> outv;
            if (!c) {
                outv.push_back({0, square});
                return outv;
            }
            for (size_t l = 0; l < c->leaves.size(); ++l) {
                Poly g = square;
                bool positive = true;
                for (const Cut &cut : c->leaves[l]) {
                    positive = clip(g, lp_of(c, v, cut.plane), cut.side, fr);
                    if (!positive) break;
                }
                if (positive) outv.push_back({l, std::move(g)});
            }
            return outv;
        };
        const auto fl = facets(cl, ll), fh = facets(ch, lh);
        const real un[3] = {axis == 0 ? 1.0L : 0.0L, axis == 1 ? 1.0L : 0.0L, axis == 2 ? 1.0L : 0.0L};
        for (const auto &[la, ga] : fl) {
            for (const auto &[lb, gb] : fh) {
                Poly g = ga;
                bool positive = true;
                if (ch) {
                    for (const Cut &cut : ch->leaves[lb]) {
                        positive = clip(g, lp_of(ch, lh, cut.plane), cut.side, fr);
                        if (!positive) break;
                    }
                }
                if (!positive) continue;
                (void)gb;
                sq.pieces.push_back({region_id(low, cl, la), region_id(high, ch, lb), static_cast<double>(area(g, un))});
            }
        }
        return sq;
    }

    void cut_squares() {
        cube_first.clear();
        int64_t next = 0;
        for (const Cube &c : cubes) {
            cube_first.push_back(next);
            next += static_cast<int64_t>(c.volume.size());
        }
        whole_first = next;
        std::vector<std::pair<size_t, int>> keys;
        for (const Cube &c : cubes) {
            for (int k = 0; k < 6; ++k) {
                const size_t j = neighbour(c.cell, k);
                keys.push_back(k % 2 ? std::make_pair(c.cell, k / 2) : std::make_pair(j, k / 2));
                if (!edge(j) && !has_cube(j)) whole.push_back(j);
            }
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        std::sort(whole.begin(), whole.end());
        whole.erase(std::unique(whole.begin(), whole.end()), whole.end());
        for (const Cube &c : cubes)
            for (size_t r = 0; r < c.volume.size(); ++r) {
                Region R;
                R.cell = c.cell;
                R.volume = c.volume[r];
                R.wall = c.wall[r];
                regions.push_back(std::move(R));
            }
        for (const size_t j : whole) {
            Region R;
            R.cell = j;
            R.volume = out->cell_volume(j);
            regions.push_back(std::move(R));
        }
        squares.resize(keys.size());
        parallel(keys.size(), threads, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; ++k) squares[k] = cut_square(keys[k].first, keys[k].second);
        });
        out->stats.squares = squares.size();
        out->stats.regions = static_cast<size_t>(whole_first);
    }

    // ---- the air ------------------------------------------------------------------------------

    void air(const double source[3]) {
        std::vector<char> node(N, 0);
        for (const size_t c : cube_cells) node[c] = 1;
        std::vector<size_t> row_start(n[1] * n[2] + 1, 0);
        std::vector<std::pair<size_t, size_t>> runs;
        for (size_t z = 0; z < n[2]; ++z)
            for (size_t y = 0; y < n[1]; ++y) {
                const size_t row = y + n[1] * z;
                row_start[row] = runs.size();
                if (y == 0 || z == 0 || y + 1 == n[1] || z + 1 == n[2]) continue;
                size_t x = 1;
                while (x + 1 < n[0]) {
                    if (node[index(x, y, z)]) {
                        ++x;
                        continue;
                    }
                    const size_t x0 = x;
                    while (x + 1 < n[0] && !node[index(x, y, z)]) ++x;
                    runs.push_back({x0, x - 1});
                }
            }
        row_start[n[1] * n[2]] = runs.size();
        const size_t R = regions.size(), base = R;
        UF uf(R + runs.size());
        auto run_of = [&](size_t cell) {
            const auto p = at(cell);
            const size_t row = p[1] + n[1] * p[2];
            const auto b = runs.begin() + static_cast<ptrdiff_t>(row_start[row]);
            const auto e = runs.begin() + static_cast<ptrdiff_t>(row_start[row + 1]);
            const auto it = std::upper_bound(b, e, p[0], [](size_t x, const auto &r) { return x < r.first; }) - 1;
            return base + static_cast<size_t>(it - runs.begin());
        };
        for (size_t z = 1; z + 1 < n[2]; ++z)
            for (size_t y = 1; y + 1 < n[1]; ++y) {
                const size_t row = y + n[1] * z;
                for (const size_t other : {row - 1, row - n[1]}) {
                    size_t i = row_start[row], j = row_start[other];
                    const size_t ie = row_start[row + 1], je = row_start[other + 1];
                    while (i < ie && j < je) {
                        if (runs[i].second < runs[j].first) ++i;
                        else if (runs[j].second < runs[i].first) ++j;
                        else {
                            uf.join(base + i, base + j);
                            if (runs[i].second < runs[j].second) ++i;
                            else ++j;
                        }
                    }
                }
            }
        for (size_t w = 0; w < whole.size(); ++w) uf.join(static_cast<size_t>(whole_first) + w, run_of(whole[w]));
        for (const Square &sq : squares)
            for (const Piece &p : sq.pieces)
                if (p.low >= 0 && p.high >= 0) uf.join(static_cast<size_t>(p.low), static_cast<size_t>(p.high));
        // The source's component.
        int64_t q[3];
        long long sc[3];
        for (int a = 0; a < 3; ++a) {
            q[a] = quantize(source[a], a);
            const auto it = std::upper_bound(G[a].begin(), G[a].end(), q[a]);
            sc[a] = static_cast<long long>(it - G[a].begin()) - 1;
        }
        bool inside = true;
        for (int a = 0; a < 3; ++a) inside = inside && sc[a] >= 1 && sc[a] + 2 <= static_cast<long long>(n[a]);
        if (!inside) {
            throw Refused("fitted walls: the source lies outside the grid's interior (in a cell on its edge or beyond): "
                          "move it inward or enlarge the grid.");
        }
        const size_t cell = index(static_cast<size_t>(sc[0]), static_cast<size_t>(sc[1]), static_cast<size_t>(sc[2]));
        size_t omega;
        if (const Cube *c = cube_at(cell)) {
            omega = uf.find(static_cast<size_t>(region_id(cell, c, static_cast<size_t>(leaf_holding(*c, q)))));
        } else {
            omega = uf.find(run_of(cell));
        }
        type.assign(N, kSolid);
        for (size_t z = 1; z + 1 < n[2]; ++z)
            for (size_t y = 1; y + 1 < n[1]; ++y) {
                const size_t row = y + n[1] * z;
                for (size_t r = row_start[row]; r < row_start[row + 1]; ++r) {
                    if (uf.find(base + r) != omega) continue;
                    for (size_t x = runs[r].first; x <= runs[r].second; ++x) type[index(x, y, z)] = kAir;
                }
            }
        for (size_t k = 0; k < cubes.size(); ++k)
            if (uf.find(static_cast<size_t>(cube_first[k] + cubes[k].centre)) == omega) type[cubes[k].cell] = kAir;
        for (size_t r = 0; r < R; ++r) regions[r].in_air = uf.find(r) == omega;
        size_t count = 0;
        for (size_t i = 0; i < N; ++i) count += type[i] != kSolid;
        if (count == 0) {
            throw Refused("fitted walls: the source lies in a closed space of the model that holds no cell centre: "
                          "move it into the room, or use a finer grid.");
        }
        if (layer > 0) {
            for (size_t i = 0; i < N; ++i) {
                if (type[i] != kAir || depth(i) > layer) continue;
                type[i] = kLayer;
                const auto p = at(i);
                for (int a = 0; a < 3; ++a) {
                    if (p[a] <= layer) out->open_sides[2 * a] = true;
                    if (p[a] + 1 + layer >= n[a]) out->open_sides[2 * a + 1] = true;
                }
            }
            for (size_t k = 0; k < cubes.size(); ++k) {
                const size_t cell2 = cubes[k].cell;
                const size_t d = depth(cell2);
                if (d > layer + 1) continue;
                bool next_to = d <= layer;
                for (int dd = 0; dd < 6 && !next_to; ++dd) next_to = type[neighbour(cell2, dd)] == kLayer;
                if (!next_to) continue;
                for (size_t r = 0; r < cubes[k].volume.size(); ++r)
                    if (regions[static_cast<size_t>(cube_first[k]) + r].in_air) {
                        throw Refused("fitted walls: the model reaches into the matched layer, or lies in a cell next to it "
                                      "that holds the air: move the model away from the grid's edge, or enlarge the margin "
                                      "round it.");
                    }
            }
        }
        owners();
    }

    void owners() {
        for (size_t k = 0; k < cubes.size(); ++k) {
            Region &g = regions[static_cast<size_t>(cube_first[k] + cubes[k].centre)];
            if (type[cubes[k].cell] == kAir) g.owner = static_cast<int64_t>(cubes[k].cell);
        }
        for (size_t w = 0; w < whole.size(); ++w)
            if (type[whole[w]] != kSolid) regions[static_cast<size_t>(whole_first) + w].owner = static_cast<int64_t>(whole[w]);
        struct Offer {
            int64_t region, owner;
            double area;
        };
        std::vector<Offer> offers;
        for (;;) {
            offers.clear();
            for (const Square &sq : squares)
                for (const Piece &p : sq.pieces) {
                    if (p.low < 0 || p.high < 0) continue;
                    const Region &A = regions[static_cast<size_t>(p.low)], &B = regions[static_cast<size_t>(p.high)];
                    auto offer = [&](const Region &to, int64_t r, const Region &by) {
                        if (to.owner >= 0 || !to.in_air || by.owner < 0) return;
                        if (type[static_cast<size_t>(by.owner)] != kAir) return;
                        offers.push_back({r, by.owner, std::max(p.open, 0.0)});
                    };
                    offer(A, p.low, B);
                    offer(B, p.high, A);
                }
            if (offers.empty()) break;
            std::stable_sort(offers.begin(), offers.end(), [](const Offer &x, const Offer &y) {
                return x.region != y.region ? x.region < y.region : x.owner < y.owner;
            });
            std::vector<std::pair<int64_t, int64_t>> given;
            for (size_t k = 0; k < offers.size();) {
                const int64_t r = offers[k].region;
                int64_t best = kNone;
                double most = -1.0;
                bool tie = false;
                while (k < offers.size() && offers[k].region == r) {
                    const int64_t o = offers[k].owner;
                    double a = 0.0;
                    while (k < offers.size() && offers[k].region == r && offers[k].owner == o) a += offers[k++].area;
                    if (a > most + 1e-12) {
                        tie = false;
                        most = a;
                        best = o;
                    } else if (std::abs(a - most) <= 1e-12) {
                        tie = true;
                    }
                }
                if (tie) out->stats.ties += 1;
                given.push_back({r, best});
            }
            for (const auto &[r, o] : given) regions[static_cast<size_t>(r)].owner = o;
        }
        for (const Region &g : regions) {
            if (g.in_air && g.owner < 0) {
                throw Refused("fitted walls: the model reaches within a cell of the matched layer: a part of the air the "
                              "source reaches is open only to the layer. Move the model away from the grid's edge, or "
                              "enlarge the margin round it.");
            }
        }
    }

    // ---- the rows -------------------------------------------------------------------------------

    int64_t owner_of(int64_t r) const { return r == kEdge ? kEdge : regions[static_cast<size_t>(r)].owner; }
    bool is_row(int64_t o) const { return o >= 0 && type[static_cast<size_t>(o)] == kAir; }

    template <typename F>
    void for_depth_one(F f) const {
        if (n[0] < 3 || n[1] < 3 || n[2] < 3) return;
        const size_t lx = n[0] - 2, ly = n[1] - 2, lz = n[2] - 2;
        for (size_t z = 1; z <= lz; ++z)
            for (size_t y = 1; y <= ly; ++y) {
                if (z == 1 || z == lz || y == 1 || y == ly) {
                    for (size_t x = 1; x <= lx; ++x) f(index(x, y, z));
                } else {
                    f(index(1, y, z));
                    if (lx > 1) f(index(lx, y, z));
                }
            }
    }

    void sum_up() {
        auto &rows = out->rows;
        for (const Region &g : regions)
            if (is_row(g.owner)) rows.push_back(g.owner);
        for_depth_one([&](size_t i) {
            if (type[i] == kAir) rows.push_back(static_cast<int64_t>(i));
        });
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
        std::vector<int32_t> row_index(N, -1);
        for (size_t r = 0; r < rows.size(); ++r) row_index[static_cast<size_t>(rows[r])] = static_cast<int32_t>(r);
        auto row = [&](int64_t cell) { return static_cast<size_t>(row_index[static_cast<size_t>(cell)]); };
        const size_t nr = rows.size(), w = materials + 1;
        auto &vol = out->air_volume, &open = out->open_area, &wall = out->wall_area, &sliver = out->sliver;
        vol.assign(nr, 0.0);
        open.assign(nr * 18, 0.0);
        wall.assign(nr * w, 0.0);
        sliver.assign(nr * 6, 0.0);
        std::vector<double> sliver_area(nr * 6, 0.0);
        auto add_wall = [&](int64_t o, const std::vector<Wall> &ws) {
            const size_t r = row(o);
            for (const Wall &x : ws) wall[r * w + (x.material < 0 ? materials : static_cast<size_t>(x.material))] += x.area;
        };
        std::vector<char> owns(nr, 0);
        for (const Region &g : regions) {
            if (!is_row(g.owner)) continue;
            const size_t r = row(g.owner), o = static_cast<size_t>(g.owner);
            owns[r] = 1;
            vol[r] += g.volume;
            add_wall(g.owner, g.wall);
            const size_t far = distance(o, g.cell);
            out->stats.farthest = std::max(out->stats.farthest, far);
            if (far > 1) out->stats.far_volume += g.volume;
            const int k = direction(o, g.cell);
            if (k >= 0 && k < 6) sliver[r * 6 + static_cast<size_t>(k)] += g.volume;
        }
        for (size_t r = 0; r < nr; ++r)
            if (!owns[r]) vol[r] = out->cell_volume(static_cast<size_t>(rows[r]));
        for (const Square &sq : squares)
            for (const Piece &p : sq.pieces) {
                const int64_t oa = owner_of(p.low), ob = owner_of(p.high);
                const double a = std::max(p.open, 0.0);
                if (oa == kEdge || ob == kEdge) {
                    const int64_t o = oa == kEdge ? ob : oa;
                    if (is_row(o)) {
                        wall[row(o) * w + materials] += a;
                        out->stats.edge_area += a;
                    }
                    continue;
                }
                if (oa < 0 || ob < 0) {
                    if ((oa >= 0) != (ob >= 0)) out->stats.dropped_area += a;
                    continue;
                }
                if (oa == ob) {
                    const Region &A = regions[static_cast<size_t>(p.low)], &B = regions[static_cast<size_t>(p.high)];
                    const size_t o = static_cast<size_t>(oa);
                    if ((A.cell == o) != (B.cell == o) && is_row(oa)) {
                        const int k = direction(o, A.cell == o ? B.cell : A.cell);
                        if (k >= 0 && k < 6) sliver_area[row(oa) * 6 + static_cast<size_t>(k)] += a;
                    }
                    continue;
                }
                const int k = direction(static_cast<size_t>(oa), static_cast<size_t>(ob));
                const bool both = is_row(oa) && is_row(ob);
                if (k < 0 || (k >= 6 && !both)) {
                    out->stats.dropped_area += a;
                    continue;
                }
                if (k >= 6) out->stats.diagonal_area += a;
                if (is_row(oa)) open[row(oa) * 18 + static_cast<size_t>(k)] += a;
                if (is_row(ob)) open[row(ob) * 18 + static_cast<size_t>(opposite_of(k))] += a;
            }
        for (size_t r = 0; r < nr; ++r) {
            const size_t o = static_cast<size_t>(rows[r]);
            if (has_cube(o)) continue;
            for (int k = 0; k < 6; ++k) {
                const size_t j = neighbour(o, k);
                if (has_cube(j)) continue;
                const double whole_face = out->face_area(o, k / 2);
                if (edge(j)) {
                    wall[r * w + materials] += whole_face;
                    out->stats.edge_area += whole_face;
                } else if (type[j] != kSolid) {
                    open[r * 18 + static_cast<size_t>(k)] += whole_face;
                }
            }
        }
        for (size_t q = 0; q < nr * 6; ++q) sliver[q] = sliver_area[q] > 1e-9 ? sliver[q] / sliver_area[q] : 0.0;
        double total = 0.0, lost = 0.0;
        auto add = [&](double x) {
            const double t = total + x;
            lost += std::abs(total) >= std::abs(x) ? (total - t) + x : (x - t) + total;
            total = t;
        };
        for (size_t r = 0; r < nr; ++r) add(vol[r]);
        for (size_t i = 0; i < N; ++i)
            if (type[i] == kAir && row_index[i] < 0) add(out->cell_volume(i));
        out->stats.air_volume = total + lost;
        thin_conduits();
    }

    // The conduit rule (cut_cells.hpp:5185-5352): an air region outside its owner's cube
    // whose open area to an edge-diagonal neighbour row is a link of the owner; the link's
    // pieces through it scaled by phi = 2 V / (A_s A_p) where that is below 1 - 1e-12, and a
    // share of its walls moved to that neighbour's row.
    void thin_conduits() {
        const auto &rows = out->rows;
        auto &open = out->open_area;
        auto &wall = out->wall_area;
        const size_t nr = rows.size(), w = materials + 1;
        auto row_of = [&](int64_t o) {
            return static_cast<size_t>(std::lower_bound(rows.begin(), rows.end(), o) - rows.begin());
        };
        auto cell_of = [&](int64_t g) { return static_cast<int64_t>(regions[static_cast<size_t>(g)].cell); };
        auto for_diagonal = [&](auto f) {
            for (const Square &sq : squares)
                for (const Piece &p : sq.pieces) {
                    const int64_t oa = owner_of(p.low), ob = owner_of(p.high);
                    if (oa == ob || !is_row(oa) || !is_row(ob)) continue;
                    const int k = direction(static_cast<size_t>(oa), static_cast<size_t>(ob));
                    if (k >= 6) f(p, oa, ob, k, std::max(p.open, 0.0));
                }
        };
        struct Link {
            double through = 0.0, phi = 1.0;
        };
        std::map<std::pair<int64_t, int>, Link> links;
        for_diagonal([&](const Piece &p, int64_t oa, int64_t ob, int k, double a) {
            if (cell_of(p.low) != oa) links[{p.low, k}].through += a;
            if (cell_of(p.high) != ob) links[{p.high, opposite_of(k)}].through += a;
        });
        out->stats.conduit_links = links.size();
        if (links.empty()) return;
        std::vector<int64_t> conduit;
        for (const auto &e : links)
            if (conduit.empty() || conduit.back() != e.first.first) conduit.push_back(e.first.first);
        auto conduit_at = [&](int64_t g) {
            return static_cast<size_t>(std::lower_bound(conduit.begin(), conduit.end(), g) - conduit.begin());
        };
        std::vector<double> contact(conduit.size(), 0.0);
        for (const Square &sq : squares)
            for (const Piece &p : sq.pieces) {
                const int64_t oa = owner_of(p.low), ob = owner_of(p.high);
                if (oa < 0 || oa != ob) continue;
                const bool ia = cell_of(p.low) == oa, ib = cell_of(p.high) == oa;
                if (ia == ib) continue;
                const int64_t g = ia ? p.high : p.low;
                const size_t x = conduit_at(g);
                if (x < conduit.size() && conduit[x] == g) contact[x] += std::max(p.open, 0.0);
            }
        bool any = false;
        for (auto &[key, L] : links) {
            const Region &g = regions[static_cast<size_t>(key.first)];
            const int kc = direction(static_cast<size_t>(g.owner), g.cell);
            const double a_s = contact[conduit_at(key.first)];
            if (kc < 0 || kc >= 6 || !(a_s > 0.0) || !(L.through > 0.0)) continue;
            const double ratio = 2.0 * g.volume / (a_s * L.through);
            L.phi = ratio >= 1.0 - 1e-12 ? 1.0 : ratio;
            out->stats.conduit_least = std::min(out->stats.conduit_least, L.phi);
            if (L.phi < 1.0) {
                out->stats.conduit_scaled += 1;
                any = true;
            }
        }
        if (!any) return;
        auto factor_of = [&](const Piece &p, int64_t oa, int64_t ob, int k) {
            double f = 1.0;
            if (cell_of(p.low) != oa) f *= links.find({p.low, k})->second.phi;
            if (cell_of(p.high) != ob) f *= links.find({p.high, opposite_of(k)})->second.phi;
            return f;
        };
        std::map<size_t, std::pair<double, double>> sums;  // (plain, scaled)
        for_diagonal([&](const Piece &p, int64_t oa, int64_t ob, int k, double) {
            if (!(factor_of(p, oa, ob, k) < 1.0)) return;
            sums[row_of(oa) * 18 + static_cast<size_t>(k)];
            sums[row_of(ob) * 18 + static_cast<size_t>(opposite_of(k))];
        });
        for_diagonal([&](const Piece &p, int64_t oa, int64_t ob, int k, double a) {
            const double f = factor_of(p, oa, ob, k);
            for (const size_t q : {row_of(oa) * 18 + static_cast<size_t>(k), row_of(ob) * 18 + static_cast<size_t>(opposite_of(k))}) {
                const auto it = sums.find(q);
                if (it == sums.end()) continue;
                it->second.first += a;
                it->second.second += a * f;
            }
        });
        for (const auto &[q, s2] : sums) open[q] = s2.second;
        for (auto it = links.begin(); it != links.end();) {
            const int64_t g = it->first.first;
            auto end = it;
            double spans = contact[conduit_at(g)];
            for (; end != links.end() && end->first.first == g; ++end) spans += end->second.through;
            const Region &R = regions[static_cast<size_t>(g)];
            const size_t o = static_cast<size_t>(R.owner), ro = row_of(R.owner);
            for (; it != end; ++it) {
                const double share = (1.0 - it->second.phi) * it->second.through / spans;
                if (!(share > 0.0)) continue;
                const size_t across = neighbour(o, it->first.second), rn = row_of(static_cast<int64_t>(across));
                if (!(rn < nr && static_cast<size_t>(rows[rn]) == across)) continue;
                for (const Wall &x : R.wall) {
                    const size_t col = x.material < 0 ? materials : static_cast<size_t>(x.material);
                    wall[ro * w + col] -= share * x.area;
                    wall[rn * w + col] += share * x.area;
                    out->stats.conduit_wall_moved += share * x.area;
                }
            }
        }
    }

    void sides() {
        for_depth_one([&](size_t i) {
            if (type[i] != kAir) return;
            const auto p = at(i);
            for (int a = 0; a < 3; ++a) {
                if (p[a] == 1) out->open_sides[2 * a] = true;
                if (p[a] + 2 == n[a]) out->open_sides[2 * a + 1] = true;
            }
        });
        for (const Square &sq : squares)
            for (const Piece &p : sq.pieces) {
                const int a = sq.axis;
                if (p.low == kEdge && p.high >= 0 && regions[static_cast<size_t>(p.high)].in_air) out->open_sides[2 * a] = true;
                if (p.high == kEdge && p.low >= 0 && regions[static_cast<size_t>(p.low)].in_air) out->open_sides[2 * a + 1] = true;
            }
    }

    // ---- ways --------------------------------------------------------------------------------

    // The orientation of d against the plane through a, b, c (exact): the sign of
    // (b - a) x (c - a) . (d - a).
    static int orient(const int64_t *a, const int64_t *b, const int64_t *c, const int64_t *d) {
        i128 e1[3], e2[3];
        int64_t e3[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = static_cast<i128>(b[k]) - a[k];
            e2[k] = static_cast<i128>(c[k]) - a[k];
            e3[k] = d[k] - a[k];
        }
        const i128 x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
        return sign(mul(x, e3[0]) + mul(y, e3[1]) + mul(z, e3[2]));
    }

    int64_t way(const int64_t *a, const int64_t *b) const {
        int64_t lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(a[k], b[k]);
            hi[k] = std::max(a[k], b[k]);
        }
        for (const FaceX &F : faces) {
            bool meets = true;
            for (int k = 0; k < 3 && meets; ++k) meets = F.hi[k] >= lo[k] && F.lo[k] <= hi[k];
            if (!meets) continue;
            const int64_t *p0 = wq[static_cast<size_t>(F.v[0])].data(), *p1 = wq[static_cast<size_t>(F.v[1])].data(),
                          *p2 = wq[static_cast<size_t>(F.v[2])].data();
            const int sa = orient(p0, p1, p2, a), sb = orient(p0, p1, p2, b);
            if (sa == 0 || sb == 0) {
                // An end on the face's plane: a crossing only where it lies inside the face.
                const int64_t *e = sa == 0 ? a : b;
                if (inside_triangle(F, e)) too_close("a source or probe lies on a face, or a cell centre next to one:");
                continue;
            }
            if (sa == sb) continue;
            const int t0 = orient(a, b, p0, p1), t1 = orient(a, b, p1, p2), t2 = orient(a, b, p2, p0);
            if (t0 == 0 || t1 == 0 || t2 == 0) {
                if ((t0 >= 0 && t1 >= 0 && t2 >= 0) || (t0 <= 0 && t1 <= 0 && t2 <= 0))
                    too_close("the way between a point and a cell centre passes through a face's edge:");
                continue;
            }
            if ((t0 > 0) == (t1 > 0) && (t1 > 0) == (t2 > 0)) return F.index;
        }
        return -1;
    }

    // Whether lattice point e, on face F's plane, lies inside F or on its border.
    bool inside_triangle(const FaceX &F, const int64_t *e) const {
        const PlaneX &P = planes[static_cast<size_t>(F.plane)];
        int a = 0;
        for (int c = 1; c < 3; ++c)
            if (mag(P.n[c]) > mag(P.n[a])) a = c;
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        int pos = 0, neg = 0;
        for (int k = 0; k < 3; ++k) {
            const auto &p = wq[static_cast<size_t>(F.v[k])], &q = wq[static_cast<size_t>(F.v[(k + 1) % 3])];
            const i128 v = (static_cast<i128>(q[b]) - p[b]) * (static_cast<i128>(e[c]) - p[c]) -
                           (static_cast<i128>(q[c]) - p[c]) * (static_cast<i128>(e[b]) - p[b]);
            pos += v > 0;
            neg += v < 0;
        }
        return pos == 0 || neg == 0;
    }

    bool on_face(const int64_t *q) const {
        for (const FaceX &F : faces) {
            bool meets = true;
            for (int k = 0; k < 3 && meets; ++k) meets = F.hi[k] >= q[k] && F.lo[k] <= q[k];
            if (!meets) continue;
            if (on_plane_exact(F.plane, q) == 0 && inside_triangle(F, q)) return true;
        }
        return false;
    }
};

// ---- the public face -----------------------------------------------------------------------

Fit::Fit(const Mesh &mesh, const int64_t *face_material, size_t materials, const GridSpec &grid, const double source[3],
         int threads)
    : impl_(std::make_unique<Impl>()) {
    Impl &m = *impl_;
    m.out = this;
    for (int a = 0; a < 3; ++a) m.n[a] = grid.n[a];
    m.sy = grid.n[0];
    m.sz = grid.n[0] * grid.n[1];
    m.N = m.sz * grid.n[2];
    m.h = grid.cell;
    m.layer = grid.layer;
    m.materials = materials;
    m.threads = std::max(threads, 1);
    if (grid.n[0] < 3 || grid.n[1] < 3 || grid.n[2] < 3) throw Refused("fitted walls: the grid needs 3 cells along each axis");
    for (size_t f = 0; f < mesh.face_count; ++f)
        for (int k = 0; k < 3; ++k)
            if (mesh.faces[3 * f + k] < 0 || static_cast<size_t>(mesh.faces[3 * f + k]) >= mesh.vertex_count)
                throw Refused("fitted walls: a face refers to a missing vertex");
    std::vector<char> given;
    m.frame(mesh, grid, given);
    m.lattice_grid(grid);
    m.mesh_on_lattice(mesh, face_material, given);
    m.faces_into_cubes();
    m.cut_cubes();
    m.cut_squares();
    m.air(source);
    m.sum_up();
    m.sides();
    kind = m.type;
}

Fit::~Fit() = default;

double Fit::cell_volume(size_t i) const {
    const auto p = impl_->at(i);
    return width[0][p[0]] * width[1][p[1]] * width[2][p[2]];
}

double Fit::face_area(size_t i, int axis) const {
    const auto p = impl_->at(i);
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    return width[b][p[b]] * width[c][p[c]];
}

Located Fit::locate(const double p[3]) const {
    const Impl &m = *impl_;
    Located L;
    for (int a = 0; a < 3; ++a) {
        L.q[a] = m.quantize(p[a], a);
        const auto &Cq = m.C[a];
        const auto it = std::upper_bound(Cq.begin(), Cq.end(), L.q[a]);
        L.j[a] = static_cast<long long>(it - Cq.begin()) - 1;
        if (L.j[a] >= 0) L.tied[a] = Cq[static_cast<size_t>(L.j[a])] == L.q[a];
        if (L.j[a] >= 0 && L.j[a] + 1 < static_cast<long long>(Cq.size())) {
            const int64_t c0 = Cq[static_cast<size_t>(L.j[a])], c1 = Cq[static_cast<size_t>(L.j[a] + 1)];
            L.lo[a] = quotient(L.q[a] - c0, c1 - c0);
            L.hi[a] = quotient(c1 - L.q[a], c1 - c0);
        }
    }
    return L;
}

void Fit::centre_q(size_t i, int64_t out[3]) const {
    const auto p = impl_->at(i);
    for (int a = 0; a < 3; ++a) out[a] = impl_->C[a][p[a]];
}

bool Fit::in_grid_box(const int64_t q[3]) const {
    for (int a = 0; a < 3; ++a)
        if (q[a] < impl_->G[a].front() || q[a] > impl_->G[a].back()) return false;
    return true;
}

int64_t Fit::way(const int64_t a[3], const int64_t b[3]) const { return impl_->way(a, b); }

bool Fit::on_face(const int64_t q[3]) const { return impl_->on_face(q); }

}  // namespace fitted
