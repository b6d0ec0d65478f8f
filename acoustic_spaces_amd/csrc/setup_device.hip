// The set-up's two passes over the whole grid, on the device: the cut links (which links
// between neighbouring cell centres a face of the mesh crosses) and the air's flood (every
// cell the source's cells reach by links not blocked). They stand for the reference's casts
// and flood in its WaveSolver constructor (cpp/include/acoustic_spaces/fdtd.hpp:446-477,
// 498-517, flood 3367-3382), and give the host set-up's results bit for bit:
//
// - Cut links: one thread a cell tests each of its links of the + half (the directions
//   whose cell offset is positive) against the faces of the triangle bins round it, with
//   the host's crossing test in the same double operations (the library is built without
//   contraction): a crossing strictly between the two centres, inside the triangle by the
//   barycentric margin. It tests every pair the host tests and some it does not; those lie
//   farther than a cell from the face's box, where no crossing and no margin case can
//   arise. A meeting inside the margins (an end within the guard of a face's plane next to
//   the face, or the crossing within the barycentric margin of an edge) is flagged for the
//   host, which refuses the geometry as it does itself.
// - Flood: rounds over the grid until none adds a cell. A round is three sweeps, along x,
//   y and z; in a sweep each thread owns a line of cells along the axis and carries the air
//   along it both ways through open links, and into its cells from open neighbours off the
//   line. The set of cells reached does not depend on the order.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <vector>

#include "setup_device.h"

namespace {

using asdev::Grid;
using asdev::Tri;

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// The host's crossing test of segment a -> b against triangle f: 0 no meeting, 1 a firm
// crossing, 2 a meeting within the barycentric margin, 3 an end within the guard of the
// plane next to the face (the host throws there).
__device__ int meet(const Tri &f, const double *a, const double *b, const Grid &g) {
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
    const double t = sa / (sa - sb);
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

// One thread a cell: its cut word, and flags for the host (bit 0: an end within the guard of
// a face's plane next to the face; bit 1: a link that met a face only within the margins and
// no face cuts firmly). The faces come from the cell's own bin: a bin lists every face whose
// box, widened by a bin each way, reaches it, and a face that a link of the cell can meet
// lies within two cells of the cell's centre, less than a bin (of three cells or more).
__global__ void cut_links_kernel(const Tri *__restrict__ tri, const uint32_t *__restrict__ bin_start,
                                 const uint32_t *__restrict__ bin_faces, const Grid g, uint16_t *__restrict__ cut,
                                 uint32_t *__restrict__ flags) {
    const uint64_t n = static_cast<uint64_t>(g.nx) * g.ny * g.nz;
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long x = static_cast<long long>(i % g.nx), y = static_cast<long long>((i / g.nx) % g.ny),
                    z = static_cast<long long>(i / (static_cast<uint64_t>(g.nx) * g.ny));
    double c[3];
    centre(g, x, y, z, c);
    const uint32_t b = static_cast<uint32_t>(x) / g.bin +
                       g.bx * (static_cast<uint32_t>(y) / g.bin + g.by * (static_cast<uint32_t>(z) / g.bin));
    uint32_t word = 0, unsure = 0, fault = 0;
    for (uint32_t q = bin_start[b]; q < bin_start[b + 1]; ++q) {
        const Tri f = tri[bin_faces[q]];
        for (uint32_t k = 0; k < g.directions; ++k) {
            if (g.plus_bit[k] < 0) continue;
            const long long tx = x + g.step[k][0], ty = y + g.step[k][1], tz = z + g.step[k][2];
            if (tx < 0 || ty < 0 || tz < 0 || tx >= g.nx || ty >= g.ny || tz >= g.nz) continue;
            double e[3];
            centre(g, tx, ty, tz, e);
            const int r = meet(f, c, e, g);
            if (r == 1) word |= 1u << g.plus_bit[k];
            if (r == 2) unsure |= 1u << g.plus_bit[k];
            if (r == 3) fault = 1;
        }
    }
    cut[i] = static_cast<uint16_t>(word);
    const uint32_t flag = fault | ((unsure & ~word) ? 2u : 0u);
    if (flag) atomicOr(&flags[0], flag);
}

__device__ __forceinline__ bool edge_cell(const Grid &g, long long x, long long y, long long z) {
    return x == 0 || y == 0 || z == 0 || x + 1 == g.nx || y + 1 == g.ny || z + 1 == g.nz;
}

// Whether the link from the cell at (x, y, z), index i, along k is cut (its own bit or its
// neighbour's, whichever direction owns the link).
__device__ __forceinline__ bool link_cut(const Grid &g, const uint16_t *cut, uint64_t i, long long nb, int k) {
    if (g.plus_bit[k] >= 0) return (cut[i] >> g.plus_bit[k]) & 1u;
    return (cut[nb] >> g.plus_bit[g.opposite[k]]) & 1u;
}

// One sweep of the flood along `axis`: each thread a line of cells along it (the lines off
// the grid's edge; x fastest across the lines where the axis is not x, so that a wave's
// loads coalesce). Forward, a cell not yet air becomes air through an open link to the
// previous cell of its line or to any neighbour off the line; backward, through an open
// link to the next cell. Open runs along the axis fill in one sweep, so a round of three
// sweeps (x, y, z) carries the air round a bend of its path in each.
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
    for (long long p = 1; p + 1 < length; ++p) {
        const uint64_t i = base + static_cast<uint64_t>(p) * s;
        if (air[i]) continue;
        bool open = p >= 2 && air[i - s] && !link_cut(g, cut, i, static_cast<long long>(i - s), minus);
        if (!open) {
            at[axis] = p;
            for (uint32_t k = 0; k < g.directions && !open; ++k) {
                if (static_cast<int>(k) == minus || static_cast<int>(k) == plus) continue;
                const long long tx = at[0] + g.step[k][0], ty = at[1] + g.step[k][1], tz = at[2] + g.step[k][2];
                if (edge_cell(g, tx, ty, tz)) continue;  // air never lies on the grid's edge
                const long long nb = static_cast<long long>(i) + g.step[k][0] + g.step[k][1] * static_cast<long long>(stride[1]) +
                                     g.step[k][2] * static_cast<long long>(stride[2]);
                open = air[nb] && !link_cut(g, cut, i, nb, static_cast<int>(k));
            }
        }
        if (open) {
            air[i] = 1;
            added = true;
        }
    }
    for (long long p = length - 3; p >= 1; --p) {
        const uint64_t i = base + static_cast<uint64_t>(p) * s;
        if (!air[i] && air[i + s] && !link_cut(g, cut, i, static_cast<long long>(i + s), plus)) {
            air[i] = 1;
            added = true;
        }
    }
    if (added) atomicOr(changed, 1u);
}

struct Buffers {
    std::vector<void *> p;
    ~Buffers() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename U>
    U *alloc(size_t count) {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(U));
        if (e != hipSuccess) throw std::runtime_error(std::string("the device set-up: hipMalloc: ") + hipGetErrorString(e));
        p.push_back(q);
        return static_cast<U *>(q);
    }
};

void ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("the device set-up: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

namespace asdev {

uint32_t cut_and_flood(const std::vector<Tri> &tri, const std::vector<uint32_t> &bin_start,
                       const std::vector<uint32_t> &bin_faces, const Grid &g, std::vector<uint16_t> &cut,
                       std::vector<uint8_t> &air, size_t &rounds, double *seconds) {
    const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
    const size_t bins = static_cast<size_t>(g.bx) * g.by * g.bz;
    if (g.bin < 3 || bin_start.size() != bins + 1 || bin_faces.size() != bin_start[bins] || air.size() != n ||
        g.nx == 0 || g.ny == 0 || g.nz == 0 || (g.directions != 6 && g.directions != 18) ||
        (g.nx + g.bin - 1) / g.bin != g.bx || (g.ny + g.bin - 1) / g.bin != g.by || (g.nz + g.bin - 1) / g.bin != g.bz)
        throw std::invalid_argument("the device set-up's grid, bins and buffers do not agree");
    if (n + 256 > (static_cast<size_t>(1) << 32)) throw std::invalid_argument("the grid is too large for the device set-up");
    for (uint32_t f : bin_faces)
        if (f >= tri.size()) throw std::invalid_argument("a bin lists a missing face");
    Buffers b;
    Tri *d_tri = b.alloc<Tri>(tri.size());
    uint32_t *d_start = b.alloc<uint32_t>(bin_start.size()), *d_faces = b.alloc<uint32_t>(bin_faces.size());
    uint16_t *d_cut = b.alloc<uint16_t>(n);
    uint8_t *d_air = b.alloc<uint8_t>(n);
    uint32_t *d_word = b.alloc<uint32_t>(1);
    if (!tri.empty()) ok(hipMemcpy(d_tri, tri.data(), tri.size() * sizeof(Tri), hipMemcpyHostToDevice), "hipMemcpy");
    ok(hipMemcpy(d_start, bin_start.data(), bin_start.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
    if (!bin_faces.empty()) {
        ok(hipMemcpy(d_faces, bin_faces.data(), bin_faces.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
    }
    const auto t0 = std::chrono::steady_clock::now();
    ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
    const uint32_t block = 256;
    hipLaunchKernelGGL(cut_links_kernel, dim3(static_cast<uint32_t>((n + block - 1) / block)), dim3(block), 0, nullptr,
                       d_tri, d_start, d_faces, g, d_cut, d_word);
    ok(hipGetLastError(), "cut_links_kernel");
    uint32_t flags = 0;
    ok(hipMemcpy(&flags, d_word, sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    rounds = 0;
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds != nullptr) seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    if (flags != 0) return flags;
    cut.resize(n);
    ok(hipMemcpy(cut.data(), d_cut, n * sizeof(uint16_t), hipMemcpyDeviceToHost), "hipMemcpy");
    ok(hipMemcpy(d_air, air.data(), n, hipMemcpyHostToDevice), "hipMemcpy");
    // A grid without cells off its edge holds no air (nor seeds).
    const bool inside = g.nx >= 3 && g.ny >= 3 && g.nz >= 3;
    const uint32_t lines[3] = {inside ? (g.ny - 2) * (g.nz - 2) : 0u, inside ? (g.nx - 2) * (g.nz - 2) : 0u,
                               inside ? (g.nx - 2) * (g.ny - 2) : 0u};
    while (inside) {
        ok(hipMemset(d_word, 0, sizeof(uint32_t)), "hipMemset");
        for (int axis = 0; axis < 3; ++axis) {
            hipLaunchKernelGGL(flood_kernel, dim3((lines[axis] + block - 1) / block), dim3(block), 0, nullptr, g, axis,
                               d_cut, d_air, d_word);
            ok(hipGetLastError(), "flood_kernel");
        }
        uint32_t changed = 0;
        ok(hipMemcpy(&changed, d_word, sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
        ++rounds;
        if (!changed) break;
        if (rounds > n) throw std::runtime_error("the device flood did not end");  // a round adds a cell or ends it
    }
    ok(hipMemcpy(air.data(), d_air, n, hipMemcpyDeviceToHost), "hipMemcpy");
    if (seconds != nullptr) seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    return 0;
}

}  // namespace asdev
