// The set-up's passes over the whole grid, on the device. They stand for the reference's
// WaveSolver constructor on walls on cell faces (cpp/include/acoustic_spaces/fdtd.hpp:
// 360-690): the casts that cut links (446-477), the flood of the air (498-517, flood
// 3367-3382), the matched layer (521-542) and the boundary cells (544-625), and give the
// host set-up's results (setup.cpp) bit for bit, refusals included. What stays on the
// host: the triangle bins, the checks of the source and the probes, and their readings
// (a few points).
//
// - Cut links: one thread a cell tests each of its links of the + half (the directions
//   whose cell offset is positive) against the faces of its own triangle bin, with the
//   host's crossing test in the same double operations (the library is built without
//   contraction): a crossing strictly between the two centres, inside the triangle by the
//   barycentric margin. A bin lists every face whose box, widened by a bin each way,
//   reaches it, and a face that a link of the cell can meet lies within two cells of the
//   cell's centre, less than a bin (of three cells or more). The cell tests every pair the
//   host tests and some it does not; those lie farther than a cell from the face's box,
//   where no crossing and no margin case can arise.
// - Flood: rounds over the grid until none adds a cell. A round is three sweeps, along x,
//   y and z; in a sweep each thread owns a line of cells along the axis and carries the air
//   along it both ways through open links, and into its cells from open neighbours off the
//   line. The set of cells reached does not depend on the order.
// - Matched layer and marks: a thread a cell. The layer's cells and the wall cells are
//   listed in grid order by a two-pass compaction (a count per block, the blocks' offsets
//   summed on the host, then each block's cells placed by wave ballots).
// - Wall cells: a thread a wall cell, twice: once to count its links and states (the host
//   sums the counts into the arrays' starts), once to write them. A contact's wall is the
//   face the link between the two centres crosses nearest its start, looked for from each
//   end in turn, the faces of the bins along the link visited in the host's order (each
//   face at the first of its bins the walk meets), with the host's ties and refusals. A
//   refusal is reported at the lowest cell that makes one, as the host's slabs report it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "setup_device.h"

namespace {

using asdev::Grid;
using asdev::Tri;

constexpr uint8_t kSolid = 0, kInterior = 1, kSpecial = 2, kLayer = 3;
constexpr double kMargin = 1.0 / 262144.0;  // the links' stability margin, 2^-18 (fdtd.hpp:17-27)
constexpr int32_t kEdgeWall = -1, kNoWall = -2;
constexpr uint32_t kBlock = 256;

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// The host's crossing test of segment a -> b against triangle f: 0 no meeting, 1 a firm
// crossing at fraction t, 2 a meeting within the barycentric margin (t set), 3 an end
// within the guard of the plane next to the face (the host refuses there).
__device__ int meet(const Tri &f, const double *a, const double *b, const Grid &g, double &t) {
    double da[3], db[3];
    for (int k = 0; k < 3; ++k) {
        da[k] = a[k] - f.v0[k];
        db[k] = b[k] - f.v0[k];
    }
    const double sa = dot3(f.n, da), sb = dot3(f.n, db);
    const auto near_box = [&](const double *p) {
        return p[0] >= f.lo[0] - g.guard && p[0] <= f.hi[0] + g.guard && p[1] >= f.lo[1] - g.guard &&
               p[1] <= f.hi[1] + g.guard && p[2] >= f.lo[2] - g.guard && p[2] <= f.hi[2] + g.guard;
    };
    if ((fabs(sa) <= g.guard && near_box(a)) || (fabs(sb) <= g.guard && near_box(b))) return 3;
    if ((sa > 0.0) == (sb > 0.0)) return 0;
    t = sa / (sa - sb);
    double p[3], h[3], s[3], q[3];
    for (int k = 0; k < 3; ++k) p[k] = a[k] + (b[k] - a[k]) * t;
    cross3(f.n, f.e2, h);
    const double det = dot3(f.e1, h);
    if (!(fabs(det) > 0.0)) return 0;
    for (int k = 0; k < 3; ++k) s[k] = p[k] - f.v0[k];
    const double u = dot3(s, h) / det;
    cross3(f.e1, f.n, q);
    const double v = dot3(s, q) / det;
    const double w = 1.0 - u - v;
    const double least = fmin(u, fmin(v, w));
    if (least > g.barycentric) return 1;
    if (least < -g.barycentric) return 0;
    return 2;
}

__device__ __forceinline__ void centre(const Grid &g, long long x, long long y, long long z, double *c) {
    c[0] = g.origin[0] + (static_cast<double>(x) + 0.5) * g.cell;
    c[1] = g.origin[1] + (static_cast<double>(y) + 0.5) * g.cell;
    c[2] = g.origin[2] + (static_cast<double>(z) + 0.5) * g.cell;
}

struct At {
    long long x, y, z;
};

__device__ __forceinline__ At at_of(const Grid &g, uint64_t i) {
    return {static_cast<long long>(i % g.nx), static_cast<long long>((i / g.nx) % g.ny),
            static_cast<long long>(i / (static_cast<uint64_t>(g.nx) * g.ny))};
}

__device__ __forceinline__ long long offset_of(const Grid &g, int k) {
    return g.step[k][0] + g.step[k][1] * static_cast<long long>(g.nx) +
           g.step[k][2] * static_cast<long long>(g.nx) * static_cast<long long>(g.ny);
}

__device__ __forceinline__ bool edge_cell(const Grid &g, long long x, long long y, long long z) {
    return x == 0 || y == 0 || z == 0 || x + 1 == g.nx || y + 1 == g.ny || z + 1 == g.nz;
}

// Whether the neighbour of (x, y, z) along k lies on the grid's edge (or past it).
__device__ __forceinline__ bool neighbour_on_edge(const Grid &g, const At &p, int k) {
    const long long to[3] = {p.x + g.step[k][0], p.y + g.step[k][1], p.z + g.step[k][2]};
    const long long size[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a)
        if (to[a] <= 0 || to[a] >= size[a] - 1) return true;
    return false;
}

// Whether the link from cell i along k is cut: its own bit, or its neighbour's, whichever
// end's direction owns the link.
__device__ __forceinline__ bool is_cut(const Grid &g, const uint16_t *cut, uint64_t i, int k) {
    if (g.plus_bit[k] >= 0) return (cut[i] >> g.plus_bit[k]) & 1u;
    const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, k));
    return (cut[j] >> g.plus_bit[g.opposite[k]]) & 1u;
}

struct Mask {
    uint32_t blocked, edge_far;
};

// A cell off the grid's edge: its blocked links and those of them whose far cell is on the
// edge.
__device__ Mask links_at(const Grid &g, const uint16_t *cut, uint64_t i, const At &p) {
    Mask m{0, 0};
    for (uint32_t k = 0; k < g.directions; ++k) {
        if (neighbour_on_edge(g, p, static_cast<int>(k))) {
            m.edge_far |= 1u << k;
            m.blocked |= 1u << k;
        } else if (is_cut(g, cut, i, static_cast<int>(k))) {
            m.blocked |= 1u << k;
        }
    }
    return m;
}

// A cell's link weights before they are averaged with its neighbours' (the host's fold):
// an open link keeps its weight; a blocked diagonal whose far cell is not on the edge hands
// its weight to its open axial sides, shared equally; every other blocked link weighs 0.
__device__ void fold(const Grid &g, const Mask &m, double *w) {
    for (int k = 0; k < 18; ++k) {
        const bool open = k < static_cast<int>(g.directions) && !((m.blocked >> k) & 1u);
        w[k] = open ? g.exact[k] : 0.0;
    }
    for (int k = 6; k < static_cast<int>(g.directions); ++k) {
        if (!((m.blocked >> k) & 1u) || ((m.edge_far >> k) & 1u)) continue;
        const bool open_a = !((m.blocked >> g.side_a[k]) & 1u), open_b = !((m.blocked >> g.side_b[k]) & 1u);
        const int sides = (open_a ? 1 : 0) + (open_b ? 1 : 0);
        if (sides == 0) continue;
        const double share = g.exact[k] / sides;
        if (open_a) w[g.side_a[k]] += share;
        if (open_b) w[g.side_b[k]] += share;
    }
}

// The largest float not above w, of w lowered by the margin (the host's lowered).
__device__ __forceinline__ float lowered(double w) {
    const double v = w * (1.0 - kMargin);
    float f = static_cast<float>(v);
    if (static_cast<double>(f) > v) f = nextafterf(f, 0.0f);
    return f;
}

// OR across a wavefront, then one atomic from its first active lane.
__device__ __forceinline__ void wave_or(uint32_t *to, uint32_t bits) {
    for (int s = 32; s >= 1; s >>= 1) bits |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(bits), s));
    if (bits != 0 && __lane_id() == static_cast<unsigned>(__ffsll(static_cast<long long>(__ballot(1))) - 1))
        atomicOr(to, bits);
}

// --- cut links and flood ------------------------------------------------------------------

__global__ void cut_links_kernel(const Tri *__restrict__ tri, const uint32_t *__restrict__ bin_start,
                                 const uint32_t *__restrict__ bin_faces, const Grid g, uint16_t *__restrict__ cut,
                                 uint32_t *__restrict__ flags) {
    const uint64_t n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const At p = at_of(g, i);
    double c[3];
    centre(g, p.x, p.y, p.z, c);
    const uint32_t b = static_cast<uint32_t>(p.x) / g.bin +
                       g.bx * (static_cast<uint32_t>(p.y) / g.bin + g.by * (static_cast<uint32_t>(p.z) / g.bin));
    uint32_t word = 0, unsure = 0, fault = 0;
    for (uint32_t q = bin_start[b]; q < bin_start[b + 1]; ++q) {
        const Tri &f = tri[bin_faces[q]];
        for (uint32_t k = 0; k < g.directions; ++k) {
            if (g.plus_bit[k] < 0) continue;
            const long long tx = p.x + g.step[k][0], ty = p.y + g.step[k][1], tz = p.z + g.step[k][2];
            if (tx < 0 || ty < 0 || tz < 0 || tx >= g.nx || ty >= g.ny || tz >= g.nz) continue;
            double e[3], t = 0.0;
            centre(g, tx, ty, tz, e);
            const int r = meet(f, c, e, g, t);
            if (r == 1) word |= 1u << g.plus_bit[k];
            if (r == 2) unsure |= 1u << g.plus_bit[k];
            if (r == 3) fault = 1;
        }
    }
    cut[i] = static_cast<uint16_t>(word);
    const uint32_t flag = fault | ((unsure & ~word) ? 2u : 0u);
    if (flag) atomicOr(&flags[0], flag);
}

// One sweep of the flood along `axis`: each thread a line of cells along it (the lines off
// the grid's edge; x fastest across the lines where the axis is not x, so that a wave's
// loads coalesce). Forward, a cell not yet air becomes air through an open link to the
// previous cell of its line or to any neighbour off the line; backward, through an open
// link to the next cell.
__global__ void flood_kernel(const Grid g, int axis, const uint16_t *__restrict__ cut, uint8_t *__restrict__ air,
                             uint32_t *__restrict__ changed) {
    const uint32_t dims[3] = {g.nx, g.ny, g.nz};
    const uint64_t stride[3] = {1, g.nx, static_cast<uint64_t>(g.nx) * g.ny};
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;  // the two axes across the lines
    const uint32_t across = dims[u] - 2;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= across * (dims[v] - 2)) return;
    long long at[3];
    at[u] = 1 + r % across;
    at[v] = 1 + r / across;
    at[axis] = 0;
    const uint64_t base = static_cast<uint64_t>(at[0]) + static_cast<uint64_t>(at[1]) * stride[1] +
                          static_cast<uint64_t>(at[2]) * stride[2];
    const uint64_t s = stride[axis];
    const int minus = 2 * axis, plus = 2 * axis + 1;
    const long long length = dims[axis];
    bool added = false;
    for (long long q = 1; q + 1 < length; ++q) {
        const uint64_t i = base + static_cast<uint64_t>(q) * s;
        if (air[i]) continue;
        bool open = q >= 2 && air[i - s] && !is_cut(g, cut, i, minus);
        if (!open) {
            at[axis] = q;
            for (uint32_t k = 0; k < g.directions && !open; ++k) {
                if (static_cast<int>(k) == minus || static_cast<int>(k) == plus) continue;
                const long long tx = at[0] + g.step[k][0], ty = at[1] + g.step[k][1], tz = at[2] + g.step[k][2];
                if (edge_cell(g, tx, ty, tz)) continue;  // air never lies on the grid's edge
                const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, static_cast<int>(k)));
                open = air[j] && !is_cut(g, cut, i, static_cast<int>(k));
            }
        }
        if (open) {
            air[i] = 1;
            added = true;
        }
    }
    for (long long q = length - 3; q >= 1; --q) {
        const uint64_t i = base + static_cast<uint64_t>(q) * s;
        if (!air[i] && air[i + s] && !is_cut(g, cut, i, plus)) {
            air[i] = 1;
            added = true;
        }
    }
    if (added) atomicOr(changed, 1u);
}

// --- matched layer and marks ------------------------------------------------------------

// Every air cell within the layer's width of the grid's edge becomes a layer cell; a face
// cutting an axial link between it and a cell off the edge is refused (bit 8 of flags);
// bits 0-5 say which sides the layer reaches.
__global__ void layer_kernel(const Grid g, const uint16_t *__restrict__ cut, uint8_t *__restrict__ type,
                             uint32_t *__restrict__ flags) {
    const uint64_t n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t bits = 0;
    if (i < n && type[i] != kSolid) {
        const At p = at_of(g, i);
        const long long at[3] = {p.x, p.y, p.z}, size[3] = {g.nx, g.ny, g.nz};
        long long least = at[0];
        for (int a = 0; a < 3; ++a) least = min(least, min(at[a], size[a] - 1 - at[a]));
        if (least <= static_cast<long long>(g.layer)) {
            for (int k = 0; k < 6; ++k)
                if (!neighbour_on_edge(g, p, k) && is_cut(g, cut, i, k)) bits |= 1u << 8;
            type[i] = kLayer;
            for (int a = 0; a < 3; ++a) {
                if (at[a] <= static_cast<long long>(g.layer)) bits |= 1u << (2 * a);
                if (at[a] + 1 + static_cast<long long>(g.layer) >= size[a]) bits |= 1u << (2 * a + 1);
            }
        }
    }
    wave_or(flags, bits);
}

// Marks 1: interior cells with a blocked link. Counts the interior cells.
__global__ void mark_kernel(const Grid g, const uint16_t *__restrict__ cut, const uint8_t *__restrict__ type,
                            uint8_t *__restrict__ mark, unsigned long long *__restrict__ interior) {
    const uint64_t n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool inside = i < n && type[i] == kInterior;
    if (i < n) mark[i] = inside && links_at(g, cut, i, at_of(g, i)).blocked != 0 ? 1 : 0;
    const unsigned long long ones = __ballot(inside);
    if (__lane_id() == static_cast<unsigned>(__ffsll(static_cast<long long>(__ballot(1))) - 1) && ones != 0)
        atomicAdd(interior, static_cast<unsigned long long>(__popcll(ones)));
}

// Marks 2: an interior cell not marked, next to a marked cell across a link whose folded
// weight differs from the stencil's (averaging changes it).
__global__ void reach_kernel(const Grid g, const uint16_t *__restrict__ cut, const uint8_t *__restrict__ type,
                             const uint8_t *__restrict__ mark, uint8_t *__restrict__ reached) {
    const uint64_t n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || mark[i] != 1) return;
    double w[18];
    fold(g, links_at(g, cut, i, at_of(g, i)), w);
    for (uint32_t k = 0; k < g.directions; ++k) {
        const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, static_cast<int>(k)));
        if (w[k] != g.exact[k] && type[j] == kInterior && !mark[j]) reached[j] = 1;
    }
}

__global__ void merge_marks_kernel(uint64_t n, const uint8_t *__restrict__ reached, uint8_t *__restrict__ mark) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && reached[i]) mark[i] = 2;
}

// --- compaction in grid order -----------------------------------------------------------

struct IsLayer {
    const uint8_t *type;
    __device__ bool operator()(uint64_t i) const { return type[i] == kLayer; }
};
struct IsMarked {
    const uint8_t *mark;
    __device__ bool operator()(uint64_t i) const { return mark[i] != 0; }
};

template <typename Pred>
__global__ void count_kernel(uint64_t n, Pred pred, uint32_t *__restrict__ counts) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int c = __syncthreads_count(i < n && pred(i));
    if (threadIdx.x == 0) counts[blockIdx.x] = static_cast<uint32_t>(c);
}

template <typename Pred>
__global__ void place_kernel(uint64_t n, Pred pred, const uint64_t *__restrict__ offsets, int64_t *__restrict__ out) {
    __shared__ uint32_t waves[kBlock / 64];
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool take = i < n && pred(i);
    const unsigned long long ones = __ballot(take);
    const unsigned lane = __lane_id(), wave = threadIdx.x / 64;
    if (lane == 0) waves[wave] = static_cast<uint32_t>(__popcll(ones));
    __syncthreads();
    uint64_t at = offsets[blockIdx.x];
    for (unsigned w = 0; w < wave; ++w) at += waves[w];
    at += static_cast<uint64_t>(__popcll(ones & ((1ull << lane) - 1ull)));
    if (take) out[at] = static_cast<int64_t>(i);
}

// --- wall cells ---------------------------------------------------------------------------

struct Walls {
    const Tri *tri;
    const uint32_t *bin_start, *bin_faces;
    const uint16_t *cut;
    const uint8_t *type, *mark;
    const float *scale;  // or null
    const double *beta;
    const uint32_t *branch_start;
    const int32_t *branch_of;
    const double *branch_a;
};

// The face nearest a that the segment a -> b crosses, or -1; a refusal in `fault`. The
// faces of the bins round the segment in the host's order: bins by z, y, x; a bin's faces
// in its list; a face at the first of its bins the walk meets.
__device__ long long crossing(const Grid &g, const Walls &w, const double *a, const double *b, uint32_t &fault) {
    uint32_t from[3], to[3];
    const uint32_t dims[3] = {g.bx, g.by, g.bz};
    for (int k = 0; k < 3; ++k) {
        const double lo = fmin(a[k], b[k]), hi = fmax(a[k], b[k]);
        const double f = floor((lo - g.origin[k]) / g.cell / static_cast<double>(g.bin)) - 1.0;
        const double l = floor((hi - g.origin[k]) / g.cell / static_cast<double>(g.bin)) + 1.0;
        const double top = static_cast<double>(dims[k] - 1);
        from[k] = static_cast<uint32_t>(fmin(fmax(f, 0.0), top));
        to[k] = static_cast<uint32_t>(fmin(fmax(l, 0.0), top));
    }
    long long firm = -1;
    double firm_t = 0.0;
    double loose_t[16];
    long long loose_face[16];
    int loose = 0;
    bool twice = false;
    for (uint32_t bz = from[2]; bz <= to[2]; ++bz)
        for (uint32_t by = from[1]; by <= to[1]; ++by)
            for (uint32_t bx = from[0]; bx <= to[0]; ++bx) {
                const uint32_t c = bx + g.bx * (by + g.by * bz);
                for (uint32_t q = w.bin_start[c]; q < w.bin_start[c + 1]; ++q) {
                    const uint32_t fi = w.bin_faces[q];
                    const Tri &f = w.tri[fi];
                    if (bx != max(from[0], f.bin_from[0]) || by != max(from[1], f.bin_from[1]) ||
                        bz != max(from[2], f.bin_from[2]))
                        continue;
                    double t = 0.0;
                    const int r = meet(f, a, b, g, t);
                    if (r == 3) {
                        fault = asdev::kGuard;
                        return -1;
                    }
                    if (r == 1) {
                        if (firm < 0 || t < firm_t) {
                            if (firm >= 0 && fabs(t - firm_t) <= 1e-12 && f.material != w.tri[firm].material) twice = true;
                            firm = fi;
                            firm_t = t;
                        } else if (fabs(t - firm_t) <= 1e-12 && f.material != w.tri[firm].material) {
                            twice = true;
                        }
                    } else if (r == 2) {
                        if (loose == 16) {
                            fault = asdev::kManyEdges;
                            return -1;
                        }
                        loose_t[loose] = t;
                        loose_face[loose++] = fi;
                    }
                }
            }
    if (twice) {
        fault = asdev::kTwice;
        return -1;
    }
    for (int q = 0; q < loose; ++q) {
        if (firm < 0) {
            fault = asdev::kEdgeOnly;
            return -1;
        }
        const bool same_place =
            fabs(loose_t[q] - firm_t) <= 1e-9 && w.tri[loose_face[q]].material == w.tri[firm].material;
        const bool farther = loose_t[q] > firm_t + 1e-9;
        if (!same_place && !farther) {
            fault = asdev::kEdgeNear;
            return -1;
        }
    }
    return firm;
}

// What cell i's contact along `dir` meets: the grid's edge where the neighbour lies on it;
// else the material of the face nearest i that the link crosses, looked for from i, then
// from the neighbour; else nothing.
__device__ int32_t wall_of(const Grid &g, const Walls &w, uint64_t i, int dir, uint32_t &fault) {
    const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, dir));
    const At pj = at_of(g, j), pi = at_of(g, i);
    if (edge_cell(g, pj.x, pj.y, pj.z)) return kEdgeWall;
    double ci[3], cj[3];
    centre(g, pi.x, pi.y, pi.z, ci);
    centre(g, pj.x, pj.y, pj.z, cj);
    long long face = crossing(g, w, ci, cj, fault);
    if (fault) return kNoWall;
    if (face < 0) face = crossing(g, w, cj, ci, fault);
    if (fault || face < 0) return kNoWall;
    return static_cast<int32_t>(w.tri[face].material);
}

// A wall cell: its pair-averaged link weights, its wall contacts and the walls they meet,
// the wall term and its branch states; COUNT: the numbers of its links and states and any
// refusal; else its arrays at its starts.
template <bool COUNT>
__global__ void wall_cells_kernel(const Grid g, const Walls w, const int64_t *__restrict__ cells, uint64_t count,
                                  uint32_t *__restrict__ links, uint32_t *__restrict__ states,
                                  unsigned long long *__restrict__ refusal, uint32_t *__restrict__ flags,
                                  const uint64_t *__restrict__ link_start, const uint64_t *__restrict__ state_start,
                                  uint8_t *__restrict__ link_direction, float *__restrict__ link_weight,
                                  float *__restrict__ coef_a, float *__restrict__ coef_b, float *__restrict__ coef_inv,
                                  int32_t *__restrict__ state_branch, float *__restrict__ state_weight) {
    const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t sides = 0;
    if (s < count) {
        const uint64_t i = static_cast<uint64_t>(cells[s]);
        const At p = at_of(g, i);
        const Mask m = links_at(g, w.cut, i, p);
        double mine[18], theirs[18], weight[18];
        fold(g, m, mine);
        for (int k = 0; k < 18; ++k) weight[k] = 0.0;
        for (uint32_t k = 0; k < g.directions; ++k) {
            const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, static_cast<int>(k)));
            if (((m.edge_far >> k) & 1u) || w.type[j] == kSolid) continue;
            double toward = g.exact[k];
            if (w.type[j] == kInterior && w.mark[j]) {
                fold(g, links_at(g, w.cut, j, at_of(g, j)), theirs);
                toward = theirs[g.opposite[k]];
            }
            weight[k] = 0.5 * (mine[k] + toward);
        }
        // Contacts in direction order; a blocked diagonal meets the walls across its
        // blocked axial sides, or the wall across itself where neither is blocked.
        double gsum = 0.0;
        int32_t met[36];
        double met_weight[36];
        int kinds = 0;
        uint32_t fault = 0;
        for (uint32_t k = 0; k < g.directions && !fault; ++k) {
            if (!((m.blocked >> k) & 1u)) continue;
            int dirs[2];
            int nd = 0;
            if (k < 6) {
                dirs[nd++] = static_cast<int>(k);
            } else {
                const bool a = (m.blocked >> g.side_a[k]) & 1u, b = (m.blocked >> g.side_b[k]) & 1u;
                if (a) dirs[nd++] = g.side_a[k];
                if (b) dirs[nd++] = g.side_b[k];
                if (!a && !b) dirs[nd++] = static_cast<int>(k);
            }
            for (int c = 0; c < nd && !fault; ++c) {
                const int dir = dirs[c];
                const int32_t wall = wall_of(g, w, i, dir, fault);
                if (fault || wall == kNoWall) continue;
                if (wall == kEdgeWall) {
                    gsum += g.wall[k] * g.edge_admittance;
                    const At pj = at_of(g, static_cast<uint64_t>(static_cast<long long>(i) + offset_of(g, dir)));
                    const long long at[3] = {pj.x, pj.y, pj.z}, size[3] = {g.nx, g.ny, g.nz};
                    sides |= 1u << 6;
                    for (int a = 0; a < 3; ++a) {
                        if (at[a] == 0) sides |= 1u << (2 * a);
                        if (at[a] + 1 >= size[a]) sides |= 1u << (2 * a + 1);
                    }
                    continue;
                }
                gsum += g.wall[k] * w.beta[wall];
                int kind = 0;
                while (kind < kinds && met[kind] != wall) ++kind;
                if (kind == kinds) {
                    met[kinds] = wall;
                    met_weight[kinds++] = 0.0;
                }
                met_weight[kind] += g.wall[k];
            }
        }
        if (fault) {
            if (COUNT) atomicMin(refusal, static_cast<unsigned long long>(i) * 8ull + fault);
        } else {
            const double scale = w.scale == nullptr ? 1.0 : static_cast<double>(w.scale[i]);
            const double r = sqrt(scale);
            uint32_t nstates = 0;
            uint64_t st = COUNT ? 0 : state_start[s];
            for (int kind = 0; kind < kinds; ++kind) {
                for (uint32_t q = w.branch_start[met[kind]]; q < w.branch_start[met[kind] + 1]; ++q) {
                    const int32_t jb = w.branch_of[q];
                    gsum += met_weight[kind] * (0.5 / w.branch_a[jb]);
                    if (!COUNT) {
                        state_branch[st] = jb;
                        state_weight[st++] = static_cast<float>(r * met_weight[kind]);
                    }
                    ++nstates;
                }
            }
            double sum = 0.0;
            uint32_t nlinks = 0;
            uint64_t at = COUNT ? 0 : link_start[s];
            for (uint32_t k = 0; k < g.directions; ++k) {
                if (!(weight[k] > 0.0)) continue;
                sum += weight[k];
                if (!COUNT) {
                    link_direction[at] = static_cast<uint8_t>(k);
                    link_weight[at++] = lowered(scale * weight[k]);
                }
                ++nlinks;
            }
            if (COUNT) {
                links[s] = nlinks;
                states[s] = nstates;
            } else {
                coef_a[s] = static_cast<float>(2.0 - scale * sum * (1.0 - kMargin));
                coef_b[s] = static_cast<float>(1.0 - r * gsum);
                coef_inv[s] = static_cast<float>(1.0 / (1.0 + r * gsum));
            }
        }
    }
    if (COUNT) wave_or(flags, sides);
}

__global__ void special_type_kernel(const int64_t *__restrict__ cells, uint64_t count, uint8_t *__restrict__ type) {
    const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < count) type[cells[s]] = kSpecial;
}

// --- host side ------------------------------------------------------------------------------

void ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("the device set-up: ") + what + ": " + hipGetErrorString(e));
}

struct Buffers {
    std::vector<void *> p;
    ~Buffers() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename U>
    U *alloc(size_t count) {
        void *q = nullptr;
        ok(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(U)), "hipMalloc");
        p.push_back(q);
        return static_cast<U *>(q);
    }
    template <typename U>
    U *upload(const U *data, size_t count) {
        U *d = alloc<U>(count);
        if (count > 0) ok(hipMemcpy(d, data, count * sizeof(U), hipMemcpyHostToDevice), "hipMemcpy");
        return d;
    }
};

// Room for `spare` more values is kept, so that the host's few appends (the source's
// cells) do not copy the arrays again.
template <typename U>
void download(std::vector<U> &to, const U *from, size_t count, size_t spare = 0) {
    to.clear();
    to.reserve(count + spare);
    to.resize(count);
    if (count > 0) ok(hipMemcpy(to.data(), from, count * sizeof(U), hipMemcpyDeviceToHost), "hipMemcpy");
}

uint32_t read_word(const uint32_t *d) {
    uint32_t v = 0;
    ok(hipMemcpy(&v, d, sizeof(v), hipMemcpyDeviceToHost), "hipMemcpy");
    return v;
}

dim3 blocks_for(uint64_t items) { return dim3(static_cast<uint32_t>(std::max<uint64_t>((items + kBlock - 1) / kBlock, 1))); }

// Exclusive sums of counts (64-bit), uploaded; returns the total.
uint64_t offsets_of(Buffers &b, const uint32_t *d_counts, size_t count, uint64_t *&d_offsets,
                    std::vector<uint64_t> *keep = nullptr) {
    std::vector<uint32_t> counts;
    download(counts, d_counts, count);
    std::vector<uint64_t> at(count + 1, 0);
    for (size_t q = 0; q < count; ++q) at[q + 1] = at[q] + counts[q];
    d_offsets = b.upload(at.data(), at.size());
    const uint64_t total = at[count];
    if (keep != nullptr) *keep = std::move(at);
    return total;
}

// The cells for which pred holds, in grid order.
template <typename Pred>
uint64_t compact(Buffers &b, uint64_t n, Pred pred, int64_t *&d_out) {
    const dim3 grid = blocks_for(n);
    uint32_t *d_counts = b.alloc<uint32_t>(grid.x);
    hipLaunchKernelGGL(count_kernel<Pred>, grid, dim3(kBlock), 0, nullptr, n, pred, d_counts);
    ok(hipGetLastError(), "count_kernel");
    uint64_t *d_offsets = nullptr;
    const uint64_t total = offsets_of(b, d_counts, grid.x, d_offsets);
    d_out = b.alloc<int64_t>(total);
    hipLaunchKernelGGL(place_kernel<Pred>, grid, dim3(kBlock), 0, nullptr, n, pred, d_offsets, d_out);
    ok(hipGetLastError(), "place_kernel");
    return total;
}

double since(std::chrono::steady_clock::time_point &t) {
    const auto now = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(now - t).count();
    t = now;
    return s;
}

}  // namespace

namespace asdev {

void set_up(const Input &in, Output &out) {
    const Grid &g = in.g;
    const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
    const size_t bins = static_cast<size_t>(g.bx) * g.by * g.bz;
    if (g.bin < 3 || in.bin_start.size() != bins + 1 || in.bin_faces.size() != in.bin_start[bins] ||
        in.seeds.size() != n || g.nx == 0 || g.ny == 0 || g.nz == 0 || (g.directions != 6 && g.directions != 18) ||
        (g.nx + g.bin - 1) / g.bin != g.bx || (g.ny + g.bin - 1) / g.bin != g.by || (g.nz + g.bin - 1) / g.bin != g.bz ||
        in.branch_start.size() != in.beta.size() + 1 || in.branch_of.size() != in.branch_start.back())
        throw std::invalid_argument("the device set-up's grid, bins and buffers do not agree");
    if (n + kBlock > (static_cast<size_t>(1) << 32))
        throw std::invalid_argument("the grid is too large for the device set-up");
    for (uint32_t f : in.bin_faces)
        if (f >= in.tri.size()) throw std::invalid_argument("a bin lists a missing face");
    for (const Tri &t : in.tri) {
        if (t.material < 0 || static_cast<size_t>(t.material) >= in.beta.size())
            throw std::invalid_argument("a face's material is missing");
        if (t.bin_from[0] >= g.bx || t.bin_from[1] >= g.by || t.bin_from[2] >= g.bz)
            throw std::invalid_argument("a face's bins lie off the grid");
    }
    for (int32_t jb : in.branch_of)
        if (jb < 0 || static_cast<size_t>(jb) >= in.branch_a.size()) throw std::invalid_argument("a branch is missing");

    auto clock = std::chrono::steady_clock::now();
    Buffers b;
    const Tri *d_tri = b.upload(in.tri.data(), in.tri.size());
    const uint32_t *d_start = b.upload(in.bin_start.data(), in.bin_start.size());
    const uint32_t *d_faces = b.upload(in.bin_faces.data(), in.bin_faces.size());
    uint16_t *d_cut = b.alloc<uint16_t>(n);
    uint32_t *d_word = b.alloc<uint32_t>(1);
    ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
    hipLaunchKernelGGL(cut_links_kernel, blocks_for(n), dim3(kBlock), 0, nullptr, d_tri, d_start, d_faces, g, d_cut,
                       d_word);
    ok(hipGetLastError(), "cut_links_kernel");
    const uint32_t cut_flags = read_word(d_word);
    out.seconds[0] = since(clock);
    if (cut_flags & 1u) {
        out.fault = kGuard;
        return;
    }
    if (cut_flags & 2u) {
        out.fault = kEdgeLink;
        return;
    }

    // The air, in place over the seeds: 0 solid, 1 interior (the classes' own values).
    uint8_t *d_type = b.upload(in.seeds.data(), n);
    const bool inside = g.nx >= 3 && g.ny >= 3 && g.nz >= 3;  // else no cell lies off the edge
    const uint32_t lines[3] = {inside ? (g.ny - 2) * (g.nz - 2) : 0u, inside ? (g.nx - 2) * (g.nz - 2) : 0u,
                               inside ? (g.nx - 2) * (g.ny - 2) : 0u};
    out.rounds = 0;
    while (inside) {
        ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
        for (int axis = 0; axis < 3; ++axis) {
            hipLaunchKernelGGL(flood_kernel, blocks_for(lines[axis]), dim3(kBlock), 0, nullptr, g, axis, d_cut, d_type,
                               d_word);
            ok(hipGetLastError(), "flood_kernel");
        }
        ++out.rounds;
        if (!read_word(d_word)) break;
        if (out.rounds > n) throw std::runtime_error("the device flood did not end");  // a round adds a cell or ends it
    }
    out.seconds[1] = since(clock);

    // The matched layer.
    if (g.layer > 0) {
        ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
        hipLaunchKernelGGL(layer_kernel, blocks_for(n), dim3(kBlock), 0, nullptr, g, d_cut, d_type, d_word);
        ok(hipGetLastError(), "layer_kernel");
        const uint32_t flags = read_word(d_word);
        if (flags & (1u << 8)) {
            out.fault = kLayerReach;
            return;
        }
        out.sides = flags & 63u;
        int64_t *d_layer = nullptr;
        const uint64_t count = compact(b, n, IsLayer{d_type}, d_layer);
        download(out.layer_cells, d_layer, count);
        out.open = count > 0;
    }

    // The marks.
    uint8_t *d_mark = b.alloc<uint8_t>(n), *d_reached = b.alloc<uint8_t>(n);
    unsigned long long *d_count = b.alloc<unsigned long long>(1);
    ok(hipMemset(d_count, 0, sizeof(unsigned long long)), "hipMemset");
    ok(hipMemset(d_reached, 0, n), "hipMemset");
    hipLaunchKernelGGL(mark_kernel, blocks_for(n), dim3(kBlock), 0, nullptr, g, d_cut, d_type, d_mark, d_count);
    ok(hipGetLastError(), "mark_kernel");
    hipLaunchKernelGGL(reach_kernel, blocks_for(n), dim3(kBlock), 0, nullptr, g, d_cut, d_type, d_mark, d_reached);
    ok(hipGetLastError(), "reach_kernel");
    hipLaunchKernelGGL(merge_marks_kernel, blocks_for(n), dim3(kBlock), 0, nullptr, static_cast<uint64_t>(n), d_reached,
                       d_mark);
    ok(hipGetLastError(), "merge_marks_kernel");
    ok(hipMemcpy(&out.interior, d_count, sizeof(unsigned long long), hipMemcpyDeviceToHost), "hipMemcpy");
    int64_t *d_cells = nullptr;
    const uint64_t count = compact(b, n, IsMarked{d_mark}, d_cells);
    out.seconds[2] = since(clock);

    // The wall cells: counts, then the arrays.
    Walls w{};
    w.tri = d_tri;
    w.bin_start = d_start;
    w.bin_faces = d_faces;
    w.cut = d_cut;
    w.type = d_type;
    w.mark = d_mark;
    w.scale = in.scale == nullptr ? nullptr : b.upload(in.scale, n);
    w.beta = b.upload(in.beta.data(), in.beta.size());
    w.branch_start = b.upload(in.branch_start.data(), in.branch_start.size());
    w.branch_of = b.upload(in.branch_of.data(), in.branch_of.size());
    w.branch_a = b.upload(in.branch_a.data(), in.branch_a.size());
    uint32_t *d_links = b.alloc<uint32_t>(count), *d_states = b.alloc<uint32_t>(count);
    unsigned long long *d_refusal = b.alloc<unsigned long long>(1);
    const unsigned long long none = ~0ull;
    ok(hipMemcpy(d_refusal, &none, sizeof(none), hipMemcpyHostToDevice), "hipMemcpy");
    ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
    hipLaunchKernelGGL(wall_cells_kernel<true>, blocks_for(count), dim3(kBlock), 0, nullptr, g, w, d_cells, count,
                       d_links, d_states, d_refusal, d_word, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr);
    ok(hipGetLastError(), "wall_cells_kernel");
    unsigned long long refusal = none;
    ok(hipMemcpy(&refusal, d_refusal, sizeof(refusal), hipMemcpyDeviceToHost), "hipMemcpy");
    if (refusal != none) {
        out.fault = static_cast<uint32_t>(refusal % 8ull);
        out.fault_cell = refusal / 8ull;
        return;
    }
    const uint32_t wall_flags = read_word(d_word);
    out.sides |= wall_flags & 63u;
    out.open = out.open || (wall_flags & (1u << 6)) != 0;
    uint64_t *d_link_start = nullptr, *d_state_start = nullptr;
    const uint64_t nlinks = offsets_of(b, d_links, count, d_link_start, &out.link_start);
    const uint64_t nstates = offsets_of(b, d_states, count, d_state_start, &out.state_start);
    out.link_start.reserve(out.link_start.size() + 64);
    out.state_start.reserve(out.state_start.size() + 64);
    uint8_t *d_direction = b.alloc<uint8_t>(nlinks);
    float *d_weight = b.alloc<float>(nlinks), *d_a = b.alloc<float>(count), *d_b = b.alloc<float>(count),
          *d_inv = b.alloc<float>(count), *d_state_weight = b.alloc<float>(nstates);
    int32_t *d_branch = b.alloc<int32_t>(nstates);
    hipLaunchKernelGGL(wall_cells_kernel<false>, blocks_for(count), dim3(kBlock), 0, nullptr, g, w, d_cells, count,
                       nullptr, nullptr, nullptr, nullptr, d_link_start, d_state_start, d_direction, d_weight, d_a, d_b,
                       d_inv, d_branch, d_state_weight);
    ok(hipGetLastError(), "wall_cells_kernel");
    hipLaunchKernelGGL(special_type_kernel, blocks_for(count), dim3(kBlock), 0, nullptr, d_cells, count, d_type);
    ok(hipGetLastError(), "special_type_kernel");
    constexpr size_t kSpareCells = 64, kSpareLinks = 64 * 18;  // the source's reading: 4 x 4 x 4 cells
    download(out.cell, d_cells, count, kSpareCells);
    download(out.link_direction, d_direction, nlinks, kSpareLinks);
    download(out.link_weight, d_weight, nlinks, kSpareLinks);
    download(out.coef_a, d_a, count, kSpareCells);
    download(out.coef_b, d_b, count, kSpareCells);
    download(out.coef_inv, d_inv, count, kSpareCells);
    download(out.state_branch, d_branch, nstates);
    download(out.state_weight, d_state_weight, nstates);
    download(out.type, d_type, n);
    download(out.cut, d_cut, n);
    out.seconds[3] = since(clock);
}

}  // namespace asdev
