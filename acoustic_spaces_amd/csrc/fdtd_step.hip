// The wave solver's time step on MI355X (gfx950): one fused launch per step.
//
// The scheme is the reference's (cpp/include/acoustic_spaces/fdtd.hpp:17-75, run() at
// 786-904): a two-level pressure leapfrog, p+ = (2 - G) p + sum_k g_k p_k - p-, over the
// grid's interior cells; the special cells (boundary, fitted-wall and source cells) by
// their links, admittance-filter branch states and source; the split-field matched layer;
// and the probes, read before the update. Every update keeps the reference's operations
// and their order, compiled without contraction (-ffp-contract=off) and with
// single-precision denormals flushed (-fgpu-flush-denormals-to-zero), as the reference
// runs with FTZ/DAZ (fdtd.hpp:314-336): in float the fields equal the reference's bit for
// bit. The probes are summed in double, as the reference's record() (fdtd.hpp:3525-3534).
//
// Layout in HBM. The two pressure fields are float (or double) arrays of nx x ny x nz
// cells, x fastest, each row padded to a pitch of px cells (a multiple of the row
// alignment, 64 floats = 256 bytes by default), so that a wavefront's 64 lanes read one
// aligned row segment. The cell classes are one byte per cell. The interior is stepped by
// tiles of 64 x 4 cells of a plane, each marching tile_z planes along z; only tiles
// holding interior cells or face specials are launched, in a list made at set-up and
// ordered by die. A special whose links are axial and weigh as the interior's (a face
// special) is a 16-byte record with its coefficients and branches in small tables; the
// other specials sit in slabs, blocks of 256 in the order of the tiles holding them,
// slot-major, each chunk of six links' directions packed into one word. The matched
// layer's nine state values are nine planes of its cells (structure of arrays).
//
// Work of one step, by block: the probes, the general specials, the face specials (by
// blocks of their own, or each tile's after its column march), the interior tiles, the
// layer cells. They write disjoint cells of the next field and read only the current
// field (and their own cells of the previous one), so one launch carries them all.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "asfdtd.h"

namespace {

constexpr int kTileX = 64;   // cells of a tile along x: one wavefront
constexpr int kTileRows = 4; // rows (y) of a tile: one wavefront each
constexpr int kBlock = kTileX * kTileRows;
constexpr uint32_t kProbesPerBlock = kBlock / kTileX;  // one wavefront a probe
constexpr uint32_t kXcds = 8;  // MI355X's dies: blocks go to them round robin by index
// Moving air over plain rows (plain_carry_kernel): a block's tile of a plane, and its frame.
constexpr int kCarryX = 64, kCarryY = 16, kCarryHalo = 6;
constexpr int kFrameX = kCarryX + 2 * kCarryHalo, kFrameY = kCarryY + 2 * kCarryHalo;
constexpr int kFrame = kFrameX * kFrameY;
constexpr uint32_t kAllInterior = 1u << 31;  // a tile's flag: every cell of it is interior
constexpr uint32_t kNoTile = 0xFFFFFFFFu;    // a block with no tile
// Cell classes on the device: the layout's 0 solid, 1 interior, 2 special (here: a general
// special, stepped by its own blocks), 3 layer; and 4, a face special (stepped in its
// tile, or by blocks of face specials).
constexpr uint8_t kClassInterior = 1, kClassFace = 4;
// In the adjoint's own classes only: an interior cell that a receiver reads (see inj).
constexpr uint8_t kClassTap = 5;
constexpr uint32_t kNoTap = 0xFFFFFFFFu;

constexpr int kSteps[18][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

// A general special's links come in chunks of kLinkChunk; a chunk's directions are packed
// into one word, five bits each, in the links' order.
constexpr int kLinkChunk = 6;

// Direction k's steps as (dx + 1) | (dy + 1) << 2 | (dz + 1) << 4, six bits a direction:
// directions 0 to 9 in one word, 10 to 17 in another.
constexpr uint64_t step_codes(int first, int count) {
    uint64_t w = 0;
    for (int d = 0; d < count; ++d) {
        const int k = first + d;
        const uint64_t code = static_cast<uint64_t>((kSteps[k][0] + 1) | (kSteps[k][1] + 1) << 2 | (kSteps[k][2] + 1) << 4);
        w |= code << (6 * d);
    }
    return w;
}
constexpr uint64_t kStepsLow = step_codes(0, 10), kStepsHigh = step_codes(10, 8);

// The cell offset of link direction d (0 to 17) on a grid of row pitch px and plane sxy.
__device__ __forceinline__ int step_offset(uint32_t d, int px, int sxy) {
    const uint64_t word = d < 10 ? kStepsLow : kStepsHigh;
    const uint32_t code = static_cast<uint32_t>(word >> (6 * (d < 10 ? d : d - 10))) & 63u;
    return (static_cast<int>(code & 3u) - 1) + (static_cast<int>((code >> 2) & 3u) - 1) * px +
           (static_cast<int>((code >> 4) & 3u) - 1) * sxy;
}

// Everything one step's launch reads, by value (kernel arguments).
template <typename T>
struct StepArgs {
    // The previous pressure (read at a cell by that cell's update only), the current one,
    // and where the next goes: a's slot (one step a launch) or another (see pair_kernel).
    // No restrict on a and o, which may be one array; each cell's update reads its a
    // before it writes its o.
    const T *a;
    const T *__restrict__ b;
    T *o;
    const T *wa;  // the branch velocities, the same way
    const T *__restrict__ wb;
    T *wo;
    // interior
    const uint8_t *__restrict__ types;
    const float *__restrict__ scale;  // padded, or null
    const uint32_t *__restrict__ tiles;
    // Per tile block and row (wavefront): the lanes whose column holds an interior cell in
    // the tile's planes; the others skip the march (null: every lane marches).
    const uint64_t *__restrict__ tile_lanes;
    uint32_t tile_blocks;
    uint32_t tiles_x, tiles_y;
    int nx, ny, nz, px;
    int tile_z;
    long long sxy;
    T axial, diagonal, centre, total;
    // specials, in blocks of kBlock (by default in the order of the tiles holding them);
    // each block's links and states in a slab of (its widest count) x kBlock, slot-major
    uint32_t specials, special_blocks;
    const uint32_t *__restrict__ sp_cell;
    const uint8_t *__restrict__ sp_links;   // links of each special
    const uint8_t *__restrict__ sp_states;  // branch states of each special
    const uint32_t *__restrict__ link_base; // per block: its link slab
    const uint32_t *__restrict__ state_base;
    const uint32_t *__restrict__ link_dirs;  // per chunk of links (see kLinkChunk), at link_base / kLinkChunk
    const float *__restrict__ link_weight;
    const float *__restrict__ coef_a;
    const T *__restrict__ coef_b;      // T: double overrides for precision 64 (as_layout)
    const T *__restrict__ coef_inv;
    const float *__restrict__ source;
    const float *__restrict__ state_weight;
    const uint16_t *__restrict__ state_branch;  // per state slot: its branch
    const float *__restrict__ branch_now;        // per branch (fdtd.hpp:391-411)
    const float *__restrict__ branch_before;
    const T *__restrict__ branch_gain;
    const float *__restrict__ signal;  // device, indexed by step
    // face specials (see face_special): each tile's after its column march, or
    const uint32_t *__restrict__ face_tile;    // per tile block: its first record (and one past the last)
    uint32_t face_blocks, face_count;          // or: all records by blocks of their own
    const uint4 *__restrict__ face_record;     // cell, links | states << 6 | coefficients << 10,
                                               // where its states start, its first four states' kinds
    const uint16_t *__restrict__ face_state_kind;  // per state: its row of face_branch
    const float4 *__restrict__ face_coef;      // coef_a, coef_b, coef_inv, source
    const float4 *__restrict__ face_branch;    // now, before, gain, weight (row 255: zeros, unless face_lookup)
    int face_lookup;                           // 1: the table has more than 255 rows (see face_load)
    // matched layer
    uint32_t layer_cells, layer_blocks;
    const uint32_t *__restrict__ layer_cell;
    uint32_t px_magic, px_shift, ny_magic, ny_shift;  // division by px and by ny (see divide)
    T *__restrict__ layer_state;            // the pressure's parts: x, y, z of each layer cell
    // The face velocities: x, y, z of each padded cell, a face at the index of the cell
    // below it; read from one slot, written to the other (the slots swap each step).
    const T *__restrict__ layer_face_in;
    T *__restrict__ layer_face_out;
    const uint8_t *__restrict__ layer_edge;  // bit axis: the cell below along axis is no layer cell
    const float *__restrict__ layer_scale;  // or null
    const float *__restrict__ keep[6];
    const float *__restrict__ push[6];
    // probes
    uint32_t probes, probe_blocks;  // probe_blocks: whole multiples of the dies, ahead of the rest
    const uint32_t *__restrict__ probe_start;
    const uint32_t *__restrict__ probe_cell;
    const double *__restrict__ probe_weight;
    double *__restrict__ out;
    unsigned long long out_stride;
    unsigned long long step;  // index into signal and out
    T *__restrict__ history;  // or null: the specials' p(n+1) - p(n-1), step-major (adjoint)
    // The adjoint's gradient sums (as_fdtd_adjoint): with `delta` (D(n + 1) of each general
    // special, or null), the step computing Q(n) adds D(n + 1) Q(n + 1) and D(n + 1) W(n + 1)
    // (read as its b and wb) into acc_g and acc_gain.
    const T *__restrict__ delta;
    double *__restrict__ acc_g;
    double *__restrict__ acc_gain;
    // The adjoint's receivers: the step computing Q(n) adds inj[u], dJ/dy of step n through
    // tap u's weights (tap_values), where it writes a tap cell: tap_of[cell] at cells of class
    // kClassTap, sp_tap[k] (or kNoTap) at special k.
    const uint32_t *__restrict__ tap_of;
    const uint32_t *__restrict__ sp_tap;
    const T *__restrict__ inj;
    // Interior cells stepped one by one: the second launch of two steps' (see pair_kernel)
    // interior cells that are not deep.
    uint32_t cell_blocks, cell_count;
    const uint32_t *__restrict__ cells;
    // The general specials, where few distinct ones repeat over the grid (a plane's cut
    // cells along its axes): each special's record (its cell, its kind, its first chunk of
    // link directions, links | states << 8) and the kinds' tables: coefficients, link weights
    // and chunks of directions (kind_links a kind), states' weights and branches (kind_states
    // a kind). The states' velocities stay in the blocks' slabs. Null sp_rec: the slabs'
    // own arrays above. (Last, so that the other arguments keep their places.)
    const uint4 *__restrict__ sp_rec;
    const float *__restrict__ kind_coef_a;
    const T *__restrict__ kind_coef_b;
    const T *__restrict__ kind_coef_inv;
    const float *__restrict__ kind_source;
    const float *__restrict__ kind_weight;
    const uint32_t *__restrict__ kind_dirs;
    const float *__restrict__ kind_state_weight;
    const uint16_t *__restrict__ kind_state_branch;
    uint32_t kind_links, kind_states;
    // Each row's interior cells as bits, a word per kTileX cells: word (z ny + y) tiles_x +
    // x / kTileX, bit x mod kTileX (see interior_column).
    const uint64_t *__restrict__ interior_bits;
    // Float handles' kinds packed for wide loads: (coef_a, coef_b, coef_inv, source) a kind,
    // (now, before, gain, weight) a kind's state; the kinds' link weights in float4 rows.
    const float4 *__restrict__ kind_coef4;
    const float4 *__restrict__ kind_state4;
};

// A face special (see the tiles' notes): a special cell whose links are axial and weigh
// as the interior's. Its record gives its cell, its links (a mask over -x, +x, -y, +y, -z,
// +z), its coefficients (an index into a small table of coef_a, coef_b, coef_inv and
// source) and its branch states (their count, and where they start in the states' arrays;
// each state an index into a small table of now, before, gain and weight). The arithmetic
// is special_cell's, in its order. A tile's face specials are stepped after its column
// pass, by all its lanes at once, while the tile's pressures are still in the caches; or
// (on grids of many tiles, where that is faster) all of them by blocks of their own.
// A face special's loads by its record, all issued before its arithmetic (face_finish):
// the fields, its first states' velocities, and its rows of the small tables (coefficients,
// state kinds). The first four states' kinds ride in the record. Only where the kinds'
// table has more than 255 rows (face_lookup) can a kind of 255 there mean "look it up in
// face_state_kind", and only then does a lane wait for a look-up before reading the rows.
constexpr int kFaceChunk = 4;  // states held in registers; a special's others are read as it goes
template <typename T>
struct FaceLoads {
    uint4 rec;  // cell, links | states << 6 | coefficients << 10, states' start, kinds
    T here, previous, nb[6];
    T wa[kFaceChunk], wb[kFaceChunk];
    float4 coef;              // coef_a, coef_b, coef_inv, source
    float4 kind[kFaceChunk];  // now, before, gain, weight
};

template <typename T, bool LOOKUP>
__device__ __forceinline__ void face_load(const StepArgs<T> &s, const uint4 rec, FaceLoads<T> &f) {
    const int i = static_cast<int>(rec.x);
    const uint32_t sbase = rec.z, ns = (rec.y >> 6) & 15u;
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    const T *__restrict__ b = s.b;
    f.rec = rec;
    f.here = b[i];
    f.previous = s.a[i];
    f.nb[0] = b[i - 1];
    f.nb[1] = b[i + 1];
    f.nb[2] = b[i - sy];
    f.nb[3] = b[i + sy];
    f.nb[4] = b[i - sz];
    f.nb[5] = b[i + sz];
#pragma unroll
    for (int u = 0; u < kFaceChunk; ++u) {
        if (static_cast<uint32_t>(u) < ns) {
            f.wa[u] = s.wa[sbase + u];
            f.wb[u] = s.wb[sbase + u];
        }
    }
    f.coef = s.face_coef[rec.y >> 10];
    if constexpr (!LOOKUP) {
        // Every kind fits its byte; a state slot the special lacks reads row 255, a row of
        // zeros, and leaves it unused.
#pragma unroll
        for (int u = 0; u < kFaceChunk; ++u) f.kind[u] = s.face_branch[(rec.w >> (8 * u)) & 0xFFu];
    } else {
#pragma unroll
        for (int u = 0; u < kFaceChunk; ++u) {
            if (static_cast<uint32_t>(u) < ns) {
                uint32_t kid = (rec.w >> (8 * u)) & 0xFFu;
                if (kid == 0xFFu) kid = s.face_state_kind[sbase + u];
                f.kind[u] = s.face_branch[kid];
            }
        }
    }
}

// The special's update (special_cell's arithmetic, in its order): its linked neighbours
// summed in the order of the directions, its states' terms, the source, coef_inv last;
// then its states' velocities.
template <typename T>
__device__ __forceinline__ void face_finish(const StepArgs<T> &s, const FaceLoads<T> &f) {
    const uint4 rec = f.rec;
    const int i = static_cast<int>(rec.x);
    const uint32_t desc = rec.y;
    const uint32_t sbase = rec.z;
    const uint32_t links = desc & 63u, ns = (desc >> 6) & 15u;
    const float sig = s.signal[s.step];
    const float4 c = f.coef;  // coef_a, coef_b, coef_inv, source
    const T w = s.axial;
    T sum = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (links & (1u << d)) sum += w * f.nb[d];
    }
    const float4 *kind = f.kind;  // now, before, gain, weight
    T own[kFaceChunk];
    T flow = 0;
#pragma unroll
    for (int u = 0; u < kFaceChunk; ++u) {
        if (static_cast<uint32_t>(u) < ns) {
            own[u] = static_cast<T>(kind[u].x) * f.wb[u] - static_cast<T>(kind[u].y) * f.wa[u];
            flow += static_cast<T>(kind[u].w) * own[u];
        }
    }
    for (uint32_t j = kFaceChunk; j < ns; ++j) {
        const uint32_t q = sbase + j;
        const float4 k = s.face_branch[s.face_state_kind[q]];
        const T term = static_cast<T>(k.x) * s.wb[q] - static_cast<T>(k.y) * s.wa[q];
        flow += static_cast<T>(k.w) * term;
    }
    T acc = static_cast<T>(c.x) * f.here;
    acc = acc + sum;
    acc = acc - static_cast<T>(c.y) * f.previous;
    acc = acc - flow;
    acc = acc + static_cast<T>(c.w) * static_cast<T>(sig);
    const T next = acc * static_cast<T>(c.z);
    s.o[i] = next;
    const T rise = next - f.previous;
#pragma unroll
    for (int u = 0; u < kFaceChunk; ++u) {
        if (static_cast<uint32_t>(u) < ns) s.wo[sbase + u] = f.wa[u] + (rise * static_cast<T>(kind[u].z) + own[u]);
    }
    for (uint32_t j = kFaceChunk; j < ns; ++j) {
        // The term again: wa has not changed since it was taken.
        const uint32_t q = sbase + j;
        const float4 k = s.face_branch[s.face_state_kind[q]];
        const T term = static_cast<T>(k.x) * s.wb[q] - static_cast<T>(k.y) * s.wa[q];
        s.wo[q] = s.wa[q] + (rise * static_cast<T>(k.z) + term);
    }
}

template <typename T, bool LOOKUP>
__device__ __forceinline__ void face_special(const StepArgs<T> &s, uint32_t r) {
    FaceLoads<T> f;
    face_load<T, LOOKUP>(s, s.face_record[r], f);
    face_finish<T>(s, f);
}

// The face specials r0, r0 + stride, ... below r1, by the look-up variant where the handle
// needs it (a choice uniform over the launch).
template <typename T>
__device__ __forceinline__ void face_specials(const StepArgs<T> &s, uint32_t r0, uint32_t r1, uint32_t stride) {
    if (s.face_lookup) {
        for (uint32_t r = r0; r < r1; r += stride) face_special<T, true>(s, r);
    } else {
        for (uint32_t r = r0; r < r1; r += stride) face_special<T, false>(s, r);
    }
}

// Face special r alone, if it is one.
template <typename T>
__device__ __forceinline__ void face_special_at(const StepArgs<T> &s, uint32_t r) {
    if (r < s.face_count) face_specials<T>(s, r, r + 1, 1);
}

// The 7-point column of one lane: planes z0 to z1 - 1 of cell (x, y), in chunks of kUnroll
// planes. Every load of a chunk (the planes above, the four in-plane neighbours, the
// previous pressure and, unless the whole tile is interior, the classes) is issued before
// any arithmetic, so a wavefront keeps some thirty loads in flight instead of waiting
// twice a plane. Each address lies in the padded grid whatever the cell's class
// (x < nx <= px, 1 <= y <= ny - 2, 1 <= z <= nz - 2), so the loads need no guard; only
// interior cells are written.
//
// Where the tile is not all interior, a lane's cells are told by the row's interior bits
// (one word a plane for the wavefront's 64 cells: `word` is the word of the lane's row at
// plane z0, the same for the whole wavefront, `plane` the words a plane), or, with TAPS, by
// the cells' classes.
template <typename T, bool SCALED, bool ALL, bool TAPS = false>
__device__ __forceinline__ void interior_column(const StepArgs<T> &s, int i, int z0, int z1, uint32_t word = 0,
                                                uint32_t plane = 0) {
    constexpr int kUnroll = 4;
    constexpr bool kBits = !ALL && !TAPS;
    // A wavefront reads its row's word by a vector load of one address, in order with its
    // other loads: the index is kept in a vector register (an empty asm statement hides
    // that it is uniform), since a scalar load's wait would also wait on every other scalar
    // load in flight.
    uint32_t at = word;
    if constexpr (kBits) asm volatile("" : "+v"(at));
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    const uint32_t lane = threadIdx.x % kTileX;
    const T *__restrict__ b = s.b;
    T col[kUnroll + 2];  // planes z - 1, z, z + 1, ..., z + kUnroll
    col[0] = b[i - sz];
    col[1] = b[i];
    for (int z = z0; z < z1; z += kUnroll, i += kUnroll * sz, at += kUnroll * plane) {
        const int count = min(kUnroll, z1 - z);
        T xm[kUnroll], xp[kUnroll], ym[kUnroll], yp[kUnroll], prev[kUnroll];
        uint8_t cls[kUnroll];
        uint64_t bits[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (u < count) {
                const int j = i + u * sz;
                col[u + 2] = b[j + sz];
                xm[u] = b[j - 1];
                xp[u] = b[j + 1];
                ym[u] = b[j - sy];
                yp[u] = b[j + sy];
                prev[u] = s.a[j];
                if constexpr (kBits) bits[u] = s.interior_bits[at + u * plane];
                else if constexpr (!ALL) cls[u] = s.types[j];
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            bool step = ALL;
            if constexpr (kBits) step = ((bits[u] >> lane) & 1u) != 0;
            else if constexpr (!ALL) step = cls[u] == kClassInterior || (TAPS && cls[u] == kClassTap);
            if (u < count && step) {
                // The reference's order (fdtd.hpp:240-288): -x, +x, -y, +y, -z, +z.
                T sum = xm[u] + xp[u];
                sum = sum + ym[u];
                sum = sum + yp[u];
                sum = sum + col[u];
                sum = sum + col[u + 2];
                T next;
                if constexpr (SCALED) {
                    const T sc = static_cast<T>(s.scale[i + u * sz]);
                    const T c0 = static_cast<T>(2) - s.total * sc;
                    const T ax = sc * s.axial;
                    next = c0 * col[u + 1] + ax * sum;
                } else {
                    next = s.centre * col[u + 1] + s.axial * sum;
                }
                T q = next - prev[u];
                if constexpr (TAPS) {
                    if (cls[u] == kClassTap) q = q + s.inj[s.tap_of[i + u * sz]];
                }
                s.o[i + u * sz] = q;
            }
        }
        if (count == kUnroll) {
            col[0] = col[kUnroll];
            col[1] = col[kUnroll + 1];
        }
    }
}

template <typename T, bool WIDE, bool SCALED, bool TAPS = false>
__device__ __forceinline__ void interior_tile(const StepArgs<T> &s, uint32_t block) {
    const uint32_t entry = s.tiles[block];
    if (entry == kNoTile) return;
    const bool all = (entry & kAllInterior) != 0;
    const uint32_t tile = entry & ~kAllInterior;
    const uint32_t tx = tile % s.tiles_x;
    const uint32_t rest = tile / s.tiles_x;
    const uint32_t ty = rest % s.tiles_y;
    const uint32_t tz = rest / s.tiles_y;
    const int x = static_cast<int>(tx) * kTileX + static_cast<int>(threadIdx.x % kTileX);
    const int y = 1 + static_cast<int>(ty) * kTileRows + static_cast<int>(threadIdx.x / kTileX);
    bool inside = x < s.nx && y < s.ny - 1;
    if (!all && s.tile_lanes != nullptr) {
        const uint32_t row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
        inside = inside && ((s.tile_lanes[block * kTileRows + row] >> (threadIdx.x % kTileX)) & 1u);
    }
    const int z0 = 1 + static_cast<int>(tz) * s.tile_z;
    const int z1 = min(z0 + s.tile_z, s.nz - 1);
    // 32-bit indices: the padded grid holds fewer than 2^31 cells (checked at set-up).
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    int i = x + y * sy + z0 * sz;
    const T *__restrict__ b = s.b;
    if constexpr (!WIDE) {
        if (inside) {
            if (all) {
                interior_column<T, SCALED, true>(s, i, z0, z1);
            } else {
                // The row's interior bits at plane z0: the same word for the whole wavefront.
                const uint32_t row = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
                const uint32_t yr = 1 + ty * kTileRows + row;
                const uint32_t plane = static_cast<uint32_t>(s.ny) * s.tiles_x;
                const uint32_t word = (static_cast<uint32_t>(z0) * static_cast<uint32_t>(s.ny) + yr) * s.tiles_x + tx;
                interior_column<T, SCALED, false, TAPS>(s, i, z0, z1, word, plane);
            }
        }
        if (!all) {
            // The tile's face specials, by all the block's lanes, after the column march.
            // (Loading a lane's first one under the march holds some twenty more registers
            // through it, for every tile: fewer wavefronts fit, and the step is slower.)
            const uint32_t r1 = s.face_tile[block + 1];
            face_specials<T>(s, s.face_tile[block] + threadIdx.x, r1, kBlock);
        }
        return;
    }
    if (!inside) return;
    T below = b[i - sz];
    T here = b[i];
    for (int z = z0; z < z1; ++z, i += sz) {
        const T above = b[i + sz];
        const uint8_t cls = s.types[i];
        if (cls == kClassInterior || (TAPS && cls == kClassTap)) {
            // The reference's order (fdtd.hpp:240-288): -x, +x, -y, +y, -z, +z.
            T sum = b[i - 1] + b[i + 1];
            sum = sum + b[i - sy];
            sum = sum + b[i + sy];
            sum = sum + below;
            sum = sum + above;
            T next;
            if constexpr (SCALED) {
                const T sc = static_cast<T>(s.scale[i]);
                const T c0 = static_cast<T>(2) - s.total * sc;
                const T ax = sc * s.axial;
                next = c0 * here + ax * sum;
            } else {
                next = s.centre * here + s.axial * sum;
            }
            if constexpr (WIDE) {
                T edges = b[i - 1 - sy] + b[i + 1 - sy];
                edges = edges + b[i - 1 + sy];
                edges = edges + b[i + 1 + sy];
                edges = edges + b[i - 1 - sz];
                edges = edges + b[i + 1 - sz];
                edges = edges + b[i - 1 + sz];
                edges = edges + b[i + 1 + sz];
                edges = edges + b[i - sy - sz];
                edges = edges + b[i + sy - sz];
                edges = edges + b[i - sy + sz];
                edges = edges + b[i + sy + sz];
                if constexpr (SCALED) {
                    const T dg = static_cast<T>(s.scale[i]) * s.diagonal;
                    next = next + dg * edges;
                } else {
                    next = next + s.diagonal * edges;
                }
            }
            T q = next - s.a[i];
            if constexpr (TAPS) {
                if (cls == kClassTap) q = q + s.inj[s.tap_of[i]];
            }
            s.o[i] = q;
        }
        below = here;
        here = above;
    }
}

// A special cell (fdtd.hpp:860-881): its links summed in their order, its branch states'
// terms, the source, coef_inv applied last; then its states' velocities. The links and
// states sit in the block's slabs (slot-major, so a wavefront's loads of a slot coalesce).
// The loads come in two waves: first everything indexed by the special itself (its cell,
// coefficients, first chunk of link directions and weights, first chunk of states), then
// everything indexed by its cell and branches (its pressures, its links' far ends, its
// branches' coefficients); chunks of links and states past the first, rare, follow one
// by one.
//
// With KINDS, what is the special's own (coefficients, links, states' weights and branches)
// comes from its kind's rows of small tables, which stay in cache, and the special's record
// gives its cell, kind and first chunk of directions in one load: the same operands, so the
// same arithmetic in the same order.
template <typename T, bool KINDS>
__device__ __forceinline__ void special_one(const StepArgs<T> &s, uint32_t k) {
    // States taken with the first loads (the rest one by one, in the same order); three in
    // double keep the step within five wavefronts a SIMD.
    constexpr int kStateChunk = sizeof(T) == 8 ? 3 : 4;
    if (k >= s.specials) return;
    const uint32_t block = k / kBlock, t = k % kBlock;
    uint4 r{};
    if constexpr (KINDS) r = s.sp_rec[k];
    const uint32_t kd = KINDS ? r.y : 0u;
    const int i = KINDS ? static_cast<int>(r.x) : static_cast<int>(s.sp_cell[k]);
    const int nl = KINDS ? static_cast<int>(r.w & 255u) : s.sp_links[k];
    const int ns = KINDS ? static_cast<int>(r.w >> 8) : s.sp_states[k];
    // Float kinds: one 16-byte load of the coefficients, float4 rows of link weights and one
    // 16-byte load a state (its branch's now, before and gain with its weight).
    constexpr bool kPacked = KINDS && sizeof(T) == 4;
    float4 cf{};
    if constexpr (kPacked) cf = s.kind_coef4[kd];
    const float ca = kPacked ? cf.x : (KINDS ? s.kind_coef_a[kd] : s.coef_a[k]);
    const float src = kPacked ? cf.w : (KINDS ? s.kind_source[kd] : s.source[k]);
    const T cb = kPacked ? static_cast<T>(cf.y) : (KINDS ? s.kind_coef_b[kd] : s.coef_b[k]);
    const T ci = kPacked ? static_cast<T>(cf.z) : (KINDS ? s.kind_coef_inv[kd] : s.coef_inv[k]);
    const float sig = s.signal[s.step];
    // Link u's weight at lb + u link_step, chunk c's directions at db + c link_step; state
    // u's weight and branch at kb + u (the kind's) or its slot (the block's slab).
    const uint32_t link_step = KINDS ? 1u : kBlock;
    const uint32_t lb = KINDS ? kd * s.kind_links : s.link_base[block] + t;
    const uint32_t db = KINDS ? kd * s.kind_links / kLinkChunk : s.link_base[block] / kLinkChunk + t;
    const uint32_t kb = KINDS ? kd * s.kind_states : 0u;
    const uint32_t sb = s.state_base[block] + t;
    const float *__restrict__ link_weight = KINDS ? s.kind_weight : s.link_weight;
    const uint32_t *__restrict__ link_dirs = KINDS ? s.kind_dirs : s.link_dirs;
    const float *__restrict__ state_weight = KINDS ? s.kind_state_weight : s.state_weight;
    const uint16_t *__restrict__ state_branch = KINDS ? s.kind_state_branch : s.state_branch;
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    const T *__restrict__ b = s.b;
    int32_t off[kLinkChunk];
    float w[kLinkChunk];
    const uint32_t dirs = KINDS ? r.z : (nl > 0 ? link_dirs[db] : 0u);
    if constexpr (kPacked) {
        const float4 *__restrict__ rows = reinterpret_cast<const float4 *>(s.kind_weight + lb);
        float4 w0{}, w1{};
        if (nl > 0) w0 = rows[0];
        if (nl > 4) w1 = rows[1];
        w[0] = w0.x;
        w[1] = w0.y;
        w[2] = w0.z;
        w[3] = w0.w;
        w[4] = w1.x;
        w[5] = w1.y;
    }
#pragma unroll
    for (int u = 0; u < kLinkChunk; ++u) {
        if (u < nl) {
            off[u] = step_offset((dirs >> (5 * u)) & 31u, sy, sz);
            if constexpr (!kPacked) w[u] = link_weight[lb + u * link_step];
        }
    }
    float sw[kStateChunk], sn[kStateChunk], sbf[kStateChunk], sg[kStateChunk];
    T swa[kStateChunk], swb[kStateChunk];
    uint32_t kind[kStateChunk];
#pragma unroll
    for (int u = 0; u < kStateChunk; ++u) {
        if (u < ns) {
            const uint32_t q = sb + u * kBlock;
            const uint32_t e = KINDS ? kb + u : q;
            if constexpr (kPacked) {
                const float4 st = s.kind_state4[e];
                sn[u] = st.x;
                sbf[u] = st.y;
                sg[u] = st.z;
                sw[u] = st.w;
            } else {
                sw[u] = state_weight[e];
                kind[u] = state_branch[e];
            }
            swa[u] = s.wa[q];
            swb[u] = s.wb[q];
        }
    }
    if constexpr (!kPacked) {
#pragma unroll
        for (int u = 0; u < kStateChunk; ++u) {
            if (u < ns) {
                sn[u] = s.branch_now[kind[u]];
                sbf[u] = s.branch_before[kind[u]];
            }
        }
    }
    const T here = b[i];
    const T previous = s.a[i];
    T v[kLinkChunk];
#pragma unroll
    for (int u = 0; u < kLinkChunk; ++u) {
        if (u < nl) v[u] = b[i + off[u]];
    }
    T sum = 0;
#pragma unroll
    for (int u = 0; u < kLinkChunk; ++u) {
        if (u < nl) sum += static_cast<T>(w[u]) * v[u];
    }
    for (int c = kLinkChunk; c < nl; c += kLinkChunk) {
        const uint32_t more = link_dirs[db + (c / kLinkChunk) * link_step];
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) {
                off[u] = step_offset((more >> (5 * u)) & 31u, sy, sz);
                w[u] = link_weight[lb + (c + u) * link_step];
            }
        }
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) v[u] = b[i + off[u]];
        }
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) sum += static_cast<T>(w[u]) * v[u];
        }
    }
    T own[kStateChunk];
    T flow = 0;
#pragma unroll
    for (int u = 0; u < kStateChunk; ++u) {
        if (u < ns) {
            own[u] = static_cast<T>(sn[u]) * swb[u] - static_cast<T>(sbf[u]) * swa[u];
            flow += static_cast<T>(sw[u]) * own[u];
        }
    }
    for (int c = kStateChunk; c < ns; ++c) {
        const uint32_t q = sb + c * kBlock;
        const uint32_t e = KINDS ? kb + c : q;
        const uint32_t br = state_branch[e];
        const T term = static_cast<T>(s.branch_now[br]) * s.wb[q] - static_cast<T>(s.branch_before[br]) * s.wa[q];
        flow += static_cast<T>(state_weight[e]) * term;
    }
    T acc = static_cast<T>(ca) * here;
    acc = acc + sum;
    acc = acc - static_cast<T>(cb) * previous;
    acc = acc - flow;
    acc = acc + static_cast<T>(src) * static_cast<T>(sig);
    const T next = acc * static_cast<T>(ci);
    s.o[i] = next;
    // For the adjoint's gradient: each special's p(n+1) - p(n-1) of this step.
    if (s.history != nullptr) s.history[s.step * s.specials + k] = next - previous;
    if (ns > 0) {
        const T rise = next - previous;
#pragma unroll
        for (int u = 0; u < kStateChunk; ++u) {
            if (u < ns) {
                const uint32_t q = sb + u * kBlock;
                const T gain = kPacked ? static_cast<T>(sg[u]) : static_cast<T>(s.branch_gain[kind[u]]);
                s.wo[q] = swa[u] + (rise * gain + own[u]);
            }
        }
        for (int c = kStateChunk; c < ns; ++c) {
            // The term again: wa has not changed since it was taken.
            const uint32_t q = sb + c * kBlock;
            const uint32_t br = state_branch[KINDS ? kb + c : q];
            const T term = static_cast<T>(s.branch_now[br]) * s.wb[q] - static_cast<T>(s.branch_before[br]) * s.wa[q];
            s.wo[q] = s.wa[q] + (rise * static_cast<T>(s.branch_gain[br]) + term);
        }
    }
}

// The special blocks' specials: block b's thread t steps special b kBlock + t, from the
// kinds' tables (KINDS: the handle has them) or the blocks' slabs.
template <typename T, bool KINDS>
__device__ __forceinline__ void special_cell(const StepArgs<T> &s, uint32_t block) {
    special_one<T, KINDS>(s, block * kBlock + threadIdx.x);
}

// n / d for n < 2^31 by a multiply: magic = floor(2^32 (2^shift - d) / d) + 1 with
// shift = ceil(log2 d) (division by invariant integers, Granlund and Montgomery 1994; made
// by divisor_of on the host).
__device__ __forceinline__ uint32_t divide(uint32_t n, uint32_t magic, uint32_t shift) {
    const uint64_t t = __umulhi(n, magic);
    return static_cast<uint32_t>((t + n) >> shift);
}

// A matched-layer cell (fdtd.hpp:2242-2266): its six face velocities, then the three
// parts of its pressure, from the current field into the next.
//
// The reference keeps six velocities a cell. Where two layer cells meet, the upper cell's
// lower face and the lower cell's upper face start at zero and take the same update from
// the same operands (keep and push at the face's index, times the same two pressures, in
// the same order), so they hold the same value at every step. One copy a face is kept:
// each cell writes its upper face, and its lower one only where the cell below is no layer
// cell (layer_edge). The faces are read from one slot and written to the other, so that a
// cell reads the lower face its neighbour rewrites in the same step before it changes.
template <typename T>
__device__ __forceinline__ void layer_cell(const StepArgs<T> &s, uint32_t c) {
    const uint32_t cell = s.layer_cell[c];
    const uint32_t edge = s.layer_edge[c];
    const long long i = cell;
    // The cell's coordinates from its padded index (no table of them to read).
    const uint32_t yz = divide(cell, s.px_magic, s.px_shift);
    const uint32_t z = divide(yz, s.ny_magic, s.ny_shift);
    const uint32_t coord[3] = {cell - yz * static_cast<uint32_t>(s.px), yz - z * static_cast<uint32_t>(s.ny), z};
    const long long stride[3] = {1, s.px, s.sxy};
    const T *__restrict__ b = s.b;
    const bool scaled = s.layer_scale != nullptr;
    // Every load first, then the arithmetic, then the stores: the compiler cannot tell that
    // the stores miss the later loads, and would otherwise wait for each axis's loads
    // before issuing the next axis's.
    T lo[3], hi[3], below[3], above[3], part[3];
    float keep_lo[3], keep_hi[3], push_lo[3], push_hi[3], keep_c[3], push_c[3];
    const T bi = b[i];
    const float scale = scaled ? s.layer_scale[c] : 0.0f;
    // A cell's three parts, and the three faces at its index, lie together: one access each.
    __builtin_memcpy(above, s.layer_face_in + 3 * i, sizeof(above));
    __builtin_memcpy(part, s.layer_state + 3 * static_cast<size_t>(c), sizeof(part));
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const uint32_t x = coord[axis];
        lo[axis] = b[i - stride[axis]];
        hi[axis] = b[i + stride[axis]];
        below[axis] = s.layer_face_in[3 * (i - stride[axis]) + axis];
        keep_lo[axis] = s.keep[2 * axis + 1][x];
        keep_hi[axis] = s.keep[2 * axis + 1][x + 1];
        push_lo[axis] = s.push[2 * axis + 1][x];
        push_hi[axis] = s.push[2 * axis + 1][x + 1];
        keep_c[axis] = s.keep[2 * axis][x];
        push_c[axis] = s.push[2 * axis][x];
    }
    T total = 0;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        below[axis] = static_cast<T>(keep_lo[axis]) * below[axis] - static_cast<T>(push_lo[axis]) * (bi - lo[axis]);
        above[axis] = static_cast<T>(keep_hi[axis]) * above[axis] - static_cast<T>(push_hi[axis]) * (hi[axis] - bi);
        if (scaled) {
            const T ps = static_cast<T>(push_c[axis]) * static_cast<T>(scale);
            part[axis] = static_cast<T>(keep_c[axis]) * part[axis] - ps * (above[axis] - below[axis]);
        } else {
            part[axis] = static_cast<T>(keep_c[axis]) * part[axis] - static_cast<T>(push_c[axis]) * (above[axis] - below[axis]);
        }
        total += part[axis];
    }
    __builtin_memcpy(s.layer_face_out + 3 * i, above, sizeof(above));
    __builtin_memcpy(s.layer_state + 3 * static_cast<size_t>(c), part, sizeof(part));
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        if ((edge >> axis) & 1u) s.layer_face_out[3 * (i - stride[axis]) + axis] = below[axis];
    }
    s.o[i] = total;
}

// The probes (fdtd.hpp:3525-3534), one wavefront each: its lanes read the probe's cells
// of `field` and form the products w * p in double at once, into `products` (kBlock
// doubles of shared memory); then one lane sums them in the order of the cells, as the
// reference sums them, into out[probe * stride + index]. Every thread of the block reaches
// both barriers.
template <typename T>
__device__ __forceinline__ void probe_sums(uint32_t probes, const uint32_t *__restrict__ probe_start,
                                           const uint32_t *__restrict__ probe_cell,
                                           const double *__restrict__ probe_weight, const T *__restrict__ field,
                                           double *__restrict__ out, unsigned long long stride,
                                           unsigned long long index, uint32_t block, double *products) {
    const uint32_t wave = threadIdx.x / kTileX, lane = threadIdx.x % kTileX;
    const uint32_t p = block * (kBlock / kTileX) + wave;
    const bool active = p < probes;
    const uint32_t k0 = active ? probe_start[p] : 0, k1 = active ? probe_start[p + 1] : 0;
    // Every wave of the block runs the same number of rounds (the block's longest probe).
    uint32_t most = k1 - k0;
    for (int w = 0; w < kBlock / kTileX; ++w) {
        const uint32_t q = block * (kBlock / kTileX) + static_cast<uint32_t>(w);
        if (q < probes) most = max(most, probe_start[q + 1] - probe_start[q]);
    }
    double value = 0.0;
    for (uint32_t base = 0; base < most; base += kTileX) {
        const uint32_t k = k0 + base + lane;
        double product = 0.0;
        if (k < k1) product = probe_weight[k] * static_cast<double>(field[probe_cell[k]]);
        products[threadIdx.x] = product;
        __syncthreads();
        if (lane == 0 && k0 + base < k1) {
            const uint32_t count = min(static_cast<uint32_t>(kTileX), k1 - k0 - base);
            const double *mine = products + wave * kTileX;
            for (uint32_t l = 0; l < count; ++l) value += mine[l];
        }
        __syncthreads();
    }
    if (active && lane == 0) out[p * stride + index] = value;
}

template <typename T>
__device__ __forceinline__ void record_probes(const StepArgs<T> &s, uint32_t block) {
    __shared__ double products[kBlock];
    probe_sums<T>(s.probes, s.probe_start, s.probe_cell, s.probe_weight, s.b, s.out, s.out_stride, s.step, block,
                  products);
}

// An interior cell of the list (7 points, one speed of sound): interior_column's update.
template <typename T>
__device__ __forceinline__ void interior_cell(const StepArgs<T> &s, int i) {
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    const T *__restrict__ b = s.b;
    T sum = b[i - 1] + b[i + 1];
    sum = sum + b[i - sy];
    sum = sum + b[i + sy];
    sum = sum + b[i - sz];
    sum = sum + b[i + sz];
    const T next = s.centre * b[i] + s.axial * sum;
    s.o[i] = next - s.a[i];
}

// One step in one launch: the probes first (a serial sum each, read from the current
// field, which nothing writes), then special blocks (their gathers are the longest chains
// of loads), face specials by blocks of their own (unless they are stepped in their
// tiles), interior tiles, layer cells and listed interior cells.
template <typename T, bool WIDE, bool SCALED, bool KINDS>
__global__ __launch_bounds__(kBlock) void step_kernel(const StepArgs<T> s) {
    uint32_t block = blockIdx.x;
    if (block < s.probe_blocks) {
        record_probes<T>(s, block);
        return;
    }
    block -= s.probe_blocks;
    if (block < s.special_blocks) {
        special_cell<T, KINDS>(s, block);
        return;
    }
    block -= s.special_blocks;
    if (block < s.face_blocks) {
        const uint32_t r = block * kBlock + threadIdx.x;
        face_special_at<T>(s, r);
        return;
    }
    block -= s.face_blocks;
    if (block < s.tile_blocks) {
        interior_tile<T, WIDE, SCALED>(s, block);
        return;
    }
    block -= s.tile_blocks;
    if (block < s.layer_blocks) {
        const uint32_t c = block * kBlock + threadIdx.x;
        if (c < s.layer_cells) layer_cell<T>(s, c);
        return;
    }
    block -= s.layer_blocks;
    if constexpr (!WIDE && !SCALED) {
        if (block < s.cell_blocks) {
            const uint32_t c = block * kBlock + threadIdx.x;
            if (c < s.cell_count) interior_cell<T>(s, static_cast<int>(s.cells[c]));
        }
    }
}

// The same work as four launches, one per kind of cell (the split launch shape).
template <typename T, bool KINDS>
__global__ __launch_bounds__(kBlock) void special_kernel(const StepArgs<T> s) {
    special_cell<T, KINDS>(s, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void face_kernel(const StepArgs<T> s) {
    const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
    face_special_at<T>(s, r);
}

template <typename T, bool WIDE, bool SCALED>
__global__ __launch_bounds__(kBlock) void interior_kernel(const StepArgs<T> s) {
    interior_tile<T, WIDE, SCALED>(s, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void layer_kernel(const StepArgs<T> s) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c < s.layer_cells) layer_cell<T>(s, c);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void probe_kernel(const StepArgs<T> s) {
    record_probes<T>(s, blockIdx.x);
}

// ---- The adjoint (time-reversed) scheme ----------------------------------------------
//
// For a closed grid (no matched layer) the step is linear in the fields,
// p(n+1) = L p(n) - d p(n-1) - coef_inv sum_s weight (now w(n) - before w(n-1)) + source, with
// L the links' row operator times coef_inv, d = coef_inv coef_b; and the branch velocities
// w(n+1) = w(n-1) + gain (p(n+1) - p(n-1)) + now w(n) - before w(n-1). Its adjoint runs the
// same two-level recursion backwards, in Q = coef_inv P (P the adjoint of p) and W (of w):
//   W(n) = now W(n+1) + (1 - before) W(n+2) + weight (before Q(n+2) - now Q(n+1))
//   Q(n) = coef_inv (coef_a Q(n+1) + sum_l w_l Q(n+1)[i + l] - coef_b Q(n+2)
//                    + sum_s gain (W(n) - W(n+2)) + (C^T dJ/dy(n))_i),
// the interior's rows being the forward interior update itself (the links' weights are
// pair-averaged, so the Laplacian is symmetric, fdtd.hpp:62-67). The gradient of J with
// respect to a special's wall term g (coef_b = 1 - g, coef_inv = 1 / (1 + g)) and a state's
// gain is, with D(n) = p(n) - p(n-2) at the special's cell,
//   dJ/dg = -sum_n D(n) Q(n),   dJ/dgain = sum_n D(n) W(n).
//
// With a speed of sound per cell (scale s, a cell's speed squared over the grid's) a row's
// links carry its own cell's s (the interior's s axial, a special's lowered s w), so the
// transpose carries each neighbour's. The fields then hold R = s Q: in R the adjoint's
// interior rows are again the forward's scaled interior update, and a special's row is
//   R(n) = coef_inv (coef_a R(n+1) + sum_l w_l R(n+1)[i + l] - coef_b R(n+2)
//                    + s (sum_s gain (W(n) - W(n+2)) + (C^T dJ/dy(n))_i)),
// its states and sums reading Q = R / s.

template <typename T, bool SCALED>
__device__ __forceinline__ void adjoint_special(const StepArgs<T> &s, uint32_t block) {
    const uint32_t t = threadIdx.x;
    const uint32_t k = block * kBlock + t;
    if (k >= s.specials) return;
    const int i = static_cast<int>(s.sp_cell[k]);
    const int nl = s.sp_links[k], ns = s.sp_states[k];
    const T ca = s.coef_a[k], cb = s.coef_b[k], ci = s.coef_inv[k];
    const uint32_t lb = s.link_base[block] + t;
    const uint32_t db = s.link_base[block] / kLinkChunk + t;
    const uint32_t sb = s.state_base[block] + t;
    const int sy = s.px, sz = static_cast<int>(s.sxy);
    const T *__restrict__ b = s.b;
    const T qb = b[i], qa = s.a[i];
    T sum = 0;
    for (int c = 0; c < nl; c += kLinkChunk) {
        int32_t off[kLinkChunk];
        float w[kLinkChunk];
        T v[kLinkChunk];
        const uint32_t dirs = s.link_dirs[db + (c / kLinkChunk) * kBlock];
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) {
                off[u] = step_offset((dirs >> (5 * u)) & 31u, sy, sz);
                w[u] = s.link_weight[lb + (c + u) * kBlock];
            }
        }
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) v[u] = b[i + off[u]];
        }
#pragma unroll
        for (int u = 0; u < kLinkChunk; ++u) {
            if (c + u < nl) sum += static_cast<T>(w[u]) * v[u];
        }
    }
    // R = s Q where the grid has a speed per cell: the states and the sums read Q.
    const T sc = SCALED ? static_cast<T>(s.scale[i]) : static_cast<T>(1);
    const T qbq = SCALED ? qb / sc : qb, qaq = SCALED ? qa / sc : qa;
    // The previous adjoint step's sums: D(n + 1) Q(n + 1), D(n + 1) W(n + 1).
    const double d = s.delta != nullptr ? static_cast<double>(s.delta[k]) : 0.0;
    if (s.delta != nullptr) s.acc_g[k] += d * static_cast<double>(qbq);
    T gsum = 0;
    for (int j = 0; j < ns; ++j) {
        const uint32_t q = sb + j * kBlock;
        const uint32_t kd = s.state_branch[q];
        const T now = s.branch_now[kd], before = s.branch_before[kd];
        const T old = s.wa[q], wb = s.wb[q];
        if (s.delta != nullptr) s.acc_gain[q] += d * static_cast<double>(wb);
        const T fresh = now * wb + (static_cast<T>(1) - before) * old +
                        static_cast<T>(s.state_weight[q]) * (before * qaq - now * qbq);
        gsum += static_cast<T>(s.branch_gain[kd]) * (fresh - old);
        s.wo[q] = fresh;
    }
    if (SCALED) gsum = sc * gsum;
    T q = ci * (ca * qb + sum - cb * qa + gsum);
    if (s.sp_tap != nullptr) {
        const uint32_t u = s.sp_tap[k];
        if (u != kNoTap) q = q + s.inj[u];
    }
    s.o[i] = q;
}

template <typename T, bool WIDE, bool SCALED>
__global__ __launch_bounds__(kBlock) void adjoint_kernel(const StepArgs<T> s) {
    const uint32_t block = blockIdx.x;
    if (block < s.special_blocks) {
        adjoint_special<T, SCALED>(s, block);
        return;
    }
    interior_tile<T, WIDE, SCALED, true>(s, block - s.special_blocks);
}

// The probes' sensitivities into the adjoint field: Q[cell] += coef_inv (sum over the taps
// on the cell of weight dJ/dy(probe, n)), one thread per distinct tap cell; ci holds each
// cell's coef_inv, times its scale where the grid has a speed per cell (R = s Q).
template <typename T>
__global__ __launch_bounds__(kBlock) void tap_values(T *__restrict__ inj, const double *__restrict__ ci,
                                                    const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ probe,
                                                    const double *__restrict__ weight,
                                                    const double *__restrict__ g, unsigned long long stride,
                                                    unsigned long long first, uint32_t taps, uint32_t rows) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= taps * rows) return;
    const uint32_t u = id % taps, row = id / taps;
    double sum = 0.0;
    for (uint32_t t = start[u]; t < start[u + 1]; ++t) sum += weight[t] * g[probe[t] * stride + first + row];
    inj[id] = static_cast<T>(ci[u] * sum);
}

// The gradient's sums: each general special's D(n) Q(n) and each of its states' D(n) W(n)
// (Q = R / s where the grid has a speed per cell).
template <typename T, bool SCALED>
__global__ __launch_bounds__(kBlock) void accumulate_kernel(const StepArgs<T> s, const T *__restrict__ q,
                                                           const T *__restrict__ w, const T *__restrict__ delta,
                                                           double *__restrict__ acc_g, double *__restrict__ acc_gain) {
    const uint32_t block = blockIdx.x, t = threadIdx.x;
    const uint32_t k = block * kBlock + t;
    if (k >= s.specials) return;
    const double d = static_cast<double>(delta[k]);
    const uint32_t cell = s.sp_cell[k];
    const T qk = SCALED ? q[cell] / static_cast<T>(s.scale[cell]) : q[cell];
    acc_g[k] += d * static_cast<double>(qk);
    const uint32_t sb = s.state_base[block] + t;
    for (int j = 0; j < s.sp_states[k]; ++j) {
        const uint32_t slot = sb + j * kBlock;
        acc_gain[slot] += d * static_cast<double>(w[slot]);
    }
}

// ---- The adjoint of the matched layer --------------------------------------------------
//
// The layer (layer_cell) steps each cell's three parts and the velocities of its six faces
// from the pressures before the step:
//   lo_a(n+1) = kf[x] lo_a(n) - pf[x] (p(n) - p_below(n)),  hi_a likewise at x + 1 with
//   (p_above(n) - p(n));  part_a(n+1) = kc[x] part_a(n) - pc[x] (hi_a(n+1) - lo_a(n+1));
//   p(n+1) = part_x + part_y + part_z
// (pc times the cell's scale where the grid has a speed per cell; a face shared by two layer
// cells is one variable of each, of equal values). Its adjoint keeps, per layer cell, the
// adjoints of its parts (LP) and of its six face velocities (LF), apart from the fields:
// the fields' layer entries stay 0, so the rows off the layer read nothing there. Each
// adjoint step, after adjoint_kernel:
//   layer_pressure: the adjoint of the cell's p(n), P = taps + sum over the faces that read
//     p(n) (its own six and its layer neighbours' facing it) of their coefficient times LF(n+1)
//     + sum over the cells off the layer whose rows read it (alpha, the row's coefficient on
//     it, in the fields' variables) of alpha times their field at n+1; and into each such
//     cell's field at n, its face's coefficient on that cell's p times LF(n+1) (mult: in the
//     field's variables); then LP(n) = kc LP(n+1) + P;
//   layer_faces: LF_lo(n) = kf[x] LF_lo(n+1) + pc LP(n), LF_hi(n) = kf[x+1] LF_hi(n+1) - pc LP(n).
template <typename T>
struct LayerAdjoint {
    uint32_t cells;
    const uint32_t *__restrict__ cell;    // padded index of each layer cell
    const uint32_t *__restrict__ index;   // per padded cell: its layer index, or kNoTap
    const float *__restrict__ keep[6];
    const float *__restrict__ push[6];
    const float *__restrict__ scale;      // per layer cell, or null
    int px, ny;
    long long sxy;
    T *__restrict__ part;                 // 3 per layer cell
    T *__restrict__ face;                 // 6 per layer cell: lo x, hi x, lo y, hi y, lo z, hi z
    const uint32_t *__restrict__ couple_start;  // per layer cell, into the couplings
    const uint8_t *__restrict__ couple_dir;     // 0..5: the direction from the layer cell
    const uint32_t *__restrict__ couple_cell;
    const T *__restrict__ couple_alpha;
    const T *__restrict__ couple_mult;
    const uint32_t *__restrict__ tap;     // per layer cell: its tap, or kNoTap
    const T *__restrict__ inj;
    const T *__restrict__ b;              // the fields at n + 1
    T *o;                                 // the fields at n (written before, added to here)
};

template <typename T>
__device__ __forceinline__ void layer_coords(const LayerAdjoint<T> &s, uint32_t i, uint32_t *at) {
    const uint32_t yz = i / static_cast<uint32_t>(s.px);
    at[0] = i - yz * static_cast<uint32_t>(s.px);
    at[2] = yz / static_cast<uint32_t>(s.ny);
    at[1] = yz - at[2] * static_cast<uint32_t>(s.ny);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void layer_pressure_adjoint(const LayerAdjoint<T> s) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= s.cells) return;
    const uint32_t i = s.cell[c];
    uint32_t at[3];
    layer_coords(s, i, at);
    const long long stride[3] = {1, s.px, s.sxy};
    const T *lf = s.face + 6 * static_cast<size_t>(c);
    T p = 0;
    if (s.tap != nullptr && s.tap[c] != kNoTap) p = s.inj[s.tap[c]];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T pf_lo = static_cast<T>(s.push[2 * a + 1][at[a]]), pf_hi = static_cast<T>(s.push[2 * a + 1][at[a] + 1]);
        p += pf_hi * lf[2 * a + 1] - pf_lo * lf[2 * a];
        const uint32_t up = s.index[i + stride[a]], down = s.index[i - stride[a]];
        if (up != kNoTap) p += pf_hi * s.face[6 * static_cast<size_t>(up) + 2 * a];
        if (down != kNoTap) p -= pf_lo * s.face[6 * static_cast<size_t>(down) + 2 * a + 1];
    }
    for (uint32_t q = s.couple_start[c]; q < s.couple_start[c + 1]; ++q) {
        const uint32_t j = s.couple_cell[q], dir = s.couple_dir[q], a = dir / 2;
        p += s.couple_alpha[q] * s.b[j];
        // The face between the two cells reads the other's p with this coefficient.
        const T term = (dir & 1u) ? -static_cast<T>(s.push[2 * a + 1][at[a] + 1]) * lf[2 * a + 1]
                                  : static_cast<T>(s.push[2 * a + 1][at[a]]) * lf[2 * a];
        atomicAdd(&s.o[j], s.couple_mult[q] * term);
    }
    T *lp = s.part + 3 * static_cast<size_t>(c);
#pragma unroll
    for (int a = 0; a < 3; ++a) lp[a] = static_cast<T>(s.keep[2 * a][at[a]]) * lp[a] + p;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void layer_faces_adjoint(const LayerAdjoint<T> s) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= s.cells) return;
    uint32_t at[3];
    layer_coords(s, s.cell[c], at);
    T *lf = s.face + 6 * static_cast<size_t>(c);
    const T *lp = s.part + 3 * static_cast<size_t>(c);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T pc = static_cast<T>(s.push[2 * a][at[a]]);
        if (s.scale != nullptr) pc = pc * static_cast<T>(s.scale[c]);
        lf[2 * a] = static_cast<T>(s.keep[2 * a + 1][at[a]]) * lf[2 * a] + pc * lp[a];
        lf[2 * a + 1] = static_cast<T>(s.keep[2 * a + 1][at[a] + 1]) * lf[2 * a + 1] - pc * lp[a];
    }
}

// ---- Two steps a launch: `pair_kernel` --------------------------------------------------
//
// Steps k and k + 1 as two launches instead of two steps of the one-step kernel (float, 7
// points, one speed of sound). The first, pair_kernel, steps every special and layer cell
// once (p(k + 1)) and, by boxes, every interior cell twice: p(k + 1) of all of them and
// p(k + 2) of the deep ones, interior cells whose six neighbours are interior too. The
// second launch (step_kernel, its tiles replaced by a list of cells) steps the specials,
// the layer and the other interior cells from p(k + 1) to p(k + 2). The fields take four
// slots: p(k - 1) and p(k) are read by the first launch and written by neither, p(k + 1)
// and p(k + 2) go to the other two.
//
// A box is one wavefront's: 60 columns (lanes 2 to 61 of a row of 64: x0 to x0 + 59) by 4
// rows (y0 to y0 + 3) by kPairZ planes, marched along z one plane at a time with every
// value in registers:
// - p(k) of 8 rows (y0 - 2 to y0 + 5) on three planes, z - 1, z and z + 1;
// - p(k + 1) of 6 rows (y0 - 1 to y0 + 4, lanes 1 to 62 valid) on three planes: plane z is
//   stepped from p(k) and p(k - 1); then plane z - 1 of the owned rows is stepped to
//   p(k + 2) from planes z - 2 to z of p(k + 1) and its own p(k).
// The x neighbours come from the next lanes (DPP row shifts across the wavefront), the y
// and z neighbours from registers. Values of cells that are not interior (or beyond the
// grid's edges, whose loads are clamped to it) are computed and never used: a deep cell's
// neighbours are interior. Each owned cell's p(k + 1) is written if it is interior, its
// p(k + 2) if it is deep. The last box of a column starts low enough to march whole
// planes, overlapping the one below it, which writes the same values there. Every update
// is the interior's arithmetic in its order, so the fields equal one step a launch's bit
// for bit.
constexpr int kBoxX = kTileX - 4;  // columns a box owns
constexpr int kBoxRows = 4;        // rows a box owns (a row group)
constexpr uint32_t kBoxWaves = kBlock / kTileX;

struct BoxArgs {
    const float *prev;  // p(k - 1)
    const float *cur;   // p(k)
    float *out1;        // p(k + 1)
    float *out2;        // p(k + 2)
    // Per plane, row group (rows 1 + 4 g to 4 + 4 g) and x: the four rows' bytes, bit 0 set
    // where the cell is interior, bit 1 where it is deep.
    const uint32_t *bits;
    const uint32_t *boxes;  // per wavefront of the box blocks: g + groups_y (bx + boxes_x bz), or kNoTile
    uint32_t box_blocks, boxes_x, groups_y;
    int ny, nz, px;
    uint32_t plane_bytes, row_bytes, field_bytes, bits_bytes;
    float centre, axial;
};

// Raw buffer access: a wavefront-uniform byte offset (soff) and the lane's (voff), against
// a resource whose range is the array's, so that an address past it reads 0 instead of
// faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), static_cast<short>(0), static_cast<int>(bytes),
                                            0x00020000);
}
__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buffer_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}
// Lane l's value of v from lane l - 1 (shr) or l + 1 (shl) of the wavefront; lanes 0 and
// 63 get 0. Every lane must be active.
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
}

template <int kPairZ>
__device__ __forceinline__ void pair_box(const BoxArgs &s, uint32_t block) {
    const int lane = static_cast<int>(threadIdx.x % kTileX);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTileX);
    const uint32_t id = s.boxes[block * kBoxWaves + wave];
    if (id == kNoTile) return;
    const uint32_t g = id % s.groups_y, rest = id / s.groups_y;
    const int bx = static_cast<int>(rest % s.boxes_x), bz = static_cast<int>(rest / s.boxes_x);
    const int x = bx * kBoxX - 2 + lane;
    const uint32_t xo = static_cast<uint32_t>(min(max(x, 0), s.px - 1)) * 4u;
    const int y0 = 1 + kBoxRows * static_cast<int>(g);
    const int z0 = min(1 + bz * kPairZ, s.nz - 1 - kPairZ);
    const bool owned = lane >= 2 && lane < kTileX - 2;
    const auto rc = buffer_of(s.cur, s.field_bytes), rp = buffer_of(s.prev, s.field_bytes);
    const auto r1 = buffer_of(s.out1, s.field_bytes), r2 = buffer_of(s.out2, s.field_bytes);
    const auto rb = buffer_of(s.bits, s.bits_bytes);
    uint32_t yo[kBoxRows + 4];  // rows y0 - 2 .. y0 + 5, clamped to the grid
#pragma unroll
    for (int q = 0; q < kBoxRows + 4; ++q) yo[q] = static_cast<uint32_t>(min(max(y0 - 2 + q, 0), s.ny - 1)) * s.row_bytes;
    const auto plane = [&](int z) { return static_cast<uint32_t>(min(max(z, 0), s.nz - 1)) * s.plane_bytes; };
    const uint32_t bits_row = g * static_cast<uint32_t>(s.px) * 4u;
    const uint32_t bits_plane = s.groups_y * static_cast<uint32_t>(s.px) * 4u;
    const float c = s.centre, w = s.axial;
    float P0[kBoxRows + 4], P1[kBoxRows + 4];  // p(k), planes z - 1 and z
    float Q0[kBoxRows + 2], Q1[kBoxRows + 2];  // p(k + 1), planes z - 2 and z - 1
#pragma unroll
    for (int q = 0; q < kBoxRows + 4; ++q) {
        P0[q] = buffer_load(rc, xo, plane(z0 - 2) + yo[q]);
        P1[q] = buffer_load(rc, xo, plane(z0 - 1) + yo[q]);
    }
#pragma unroll
    for (int r = 0; r < kBoxRows + 2; ++r) {
        Q0[r] = 0.0f;
        Q1[r] = 0.0f;
    }
    uint32_t bits_before = 0;
#pragma unroll
    for (int t = 0; t <= kPairZ + 1; ++t) {
        const int z = z0 - 1 + t;  // the plane of p(k + 1) this pass steps
        float P2[kBoxRows + 4], A[kBoxRows + 2];
        const uint32_t pz = plane(z), pn = plane(z + 1);
#pragma unroll
        for (int q = 0; q < kBoxRows + 4; ++q) P2[q] = buffer_load(rc, xo, pn + yo[q]);
#pragma unroll
        for (int r = 0; r < kBoxRows + 2; ++r) A[r] = buffer_load(rp, xo, pz + yo[r + 1]);
        const uint32_t bits =
            __float_as_uint(buffer_load(rb, xo, static_cast<uint32_t>(min(max(z, 0), s.nz - 1)) * bits_plane + bits_row));
        // p(k + 1) of plane z, rows y0 - 1 .. y0 + 4: the interior update (fdtd.hpp:240-288)
        // in its order, -x, +x, -y, +y, -z, +z.
        float N[kBoxRows + 2];
#pragma unroll
        for (int r = 0; r < kBoxRows + 2; ++r) {
            const int q = r + 1;
            float sum = from_left(P1[q]) + from_right(P1[q]);
            sum = sum + P1[q - 1];
            sum = sum + P1[q + 1];
            sum = sum + P0[q];
            sum = sum + P2[q];
            const float next = c * P1[q] + w * sum;
            N[r] = next - A[r];
        }
        // p(k + 2) of plane z - 1, the owned rows (on every lane: the shifts read them all).
        float V[kBoxRows];
#pragma unroll
        for (int r = 1; r <= kBoxRows; ++r) {
            float sum = from_left(Q1[r]) + from_right(Q1[r]);
            sum = sum + Q1[r - 1];
            sum = sum + Q1[r + 1];
            sum = sum + Q0[r];
            sum = sum + N[r];
            const float next = c * Q1[r] + w * sum;
            V[r - 1] = next - P0[r + 1];
        }
        if (owned && t >= 1 && t <= kPairZ) {
#pragma unroll
            for (int r = 0; r < kBoxRows; ++r) {
                if ((bits >> (8 * r)) & 1u) buffer_store(N[r + 1], r1, xo, pz + yo[r + 2]);
            }
        }
        if (owned && t >= 2) {
            const uint32_t pm = plane(z - 1);
#pragma unroll
            for (int r = 0; r < kBoxRows; ++r) {
                if ((bits_before >> (8 * r)) & 2u) buffer_store(V[r], r2, xo, pm + yo[r + 2]);
            }
        }
#pragma unroll
        for (int q = 0; q < kBoxRows + 4; ++q) {
            P0[q] = P1[q];
            P1[q] = P2[q];
        }
#pragma unroll
        for (int r = 0; r < kBoxRows + 2; ++r) {
            Q0[r] = Q1[r];
            Q1[r] = N[r];
        }
        bits_before = bits;
    }
}

// The first launch of two steps: probes (of p(k)), general specials, face specials, the
// interior's boxes, layer cells. Every role but the boxes is one step's (p(k + 1) into s.o).
template <int kPairZ>
__global__ __launch_bounds__(kBlock) void pair_kernel(const StepArgs<float> s, const BoxArgs x) {
    uint32_t block = blockIdx.x;
    if (block < s.probe_blocks) {
        record_probes<float>(s, block);
        return;
    }
    block -= s.probe_blocks;
    if (block < s.special_blocks) {
        if (s.sp_rec != nullptr)
            special_cell<float, true>(s, block);
        else
            special_cell<float, false>(s, block);
        return;
    }
    block -= s.special_blocks;
    if (block < s.face_blocks) {
        const uint32_t r = block * kBlock + threadIdx.x;
        face_special_at<float>(s, r);
        return;
    }
    block -= s.face_blocks;
    if (block < x.box_blocks) {
        pair_box<kPairZ>(x, block);
        return;
    }
    block -= x.box_blocks;
    if (block < s.layer_blocks) {
        const uint32_t c = block * kBlock + threadIdx.x;
        if (c < s.layer_cells) layer_cell<float>(s, c);
        return;
    }
}

thread_local std::string g_create_error;

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct as_fdtd {
    int precision = 32;
    int wide = 0, scaled = 0;
    int nx = 0, ny = 0, nz = 0, px = 0, tile_z = 8;
    long long sxy = 0;
    size_t n_pad = 0;
    uint32_t tiles_x = 0, tiles_y = 0;
    uint32_t tile_count = 0, specials = 0, states = 0, layer_cells = 0, probes = 0;
    uint64_t interior_cells = 0;
    double axial = 0, diagonal = 0, centre = 0, total = 0;
    unsigned long long done = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocations;
    // device arrays
    void *u0 = nullptr, *u1 = nullptr, *w0 = nullptr, *w1 = nullptr, *layer_state = nullptr;
    // The fields and branch velocities by slot: 0 and 1 always, 2 and 3 for two steps a
    // launch. p(k - 1) is in slot `prev` and p(k) in prev ^ 1 (the same for the velocities).
    void *field[4] = {}, *velocity[4] = {};
    int prev = 0;
    int device = 0;     // the device the handle's memory and stream live on
    int pairs = 0;      // 1: two steps a launch (see pair_kernel)
    int pair_z = 8;     // planes a box marches
    uint32_t box_blocks = 0, boxes_x = 0, groups_y = 0, cell_count = 0;
    uint32_t *boxes = nullptr, *box_bits = nullptr, *cells = nullptr;
    uint8_t *types = nullptr;
    float *scale = nullptr;
    // Moving air (see carry_kernel), padded like the fields: plain (flow_x, flow_y and the
    // row masks) or round (each cell's +x, +y, +z faces' wind and loops, the mask 0/1/2),
    // and the carrying's two stage fields.
    bool flowing = false, flow_round = false;
    uint8_t *flow_mask = nullptr;
    float *flow_x = nullptr, *flow_y = nullptr, *flow_wind[3] = {}, *flow_loop[3] = {};
    float *flow_t1 = nullptr, *flow_t2 = nullptr;
    // The pressure squared over the air (cells not solid) as each step's carrying takes it:
    // a sum per block of the carrying's grid over the last run's steps, and their total; the
    // air's cells and the last run's steps (mean_square).
    double *flow_squares = nullptr, *flow_total = nullptr;
    uint32_t flow_blocks = 0;
    uint64_t air_cells = 0, run_steps = 0;
    uint32_t *tiles = nullptr;
    uint64_t *tile_lanes = nullptr;  // per tile block and row: its lanes that march (see interior_tile)
    uint64_t *interior_bits = nullptr;  // each row's interior cells (see StepArgs::interior_bits)
    uint32_t *sp_cell = nullptr, *link_base = nullptr, *state_base = nullptr;
    uint8_t *sp_links = nullptr, *sp_states = nullptr;
    uint32_t *link_dirs = nullptr;
    float *link_weight = nullptr, *coef_a = nullptr, *source = nullptr;
    void *coef_b = nullptr, *coef_inv = nullptr, *branch_gain = nullptr;  // float or double, as the precision
    float *state_weight = nullptr, *branch_now = nullptr, *branch_before = nullptr;
    uint16_t *state_branch = nullptr;
    // The general specials' kinds (see StepArgs::sp_rec), or null sp_rec.
    uint4 *sp_rec = nullptr;
    float *kind_coef_a = nullptr, *kind_source = nullptr, *kind_weight = nullptr, *kind_state_weight = nullptr;
    void *kind_coef_b = nullptr, *kind_coef_inv = nullptr;  // float or double, as the precision
    uint32_t *kind_dirs = nullptr;
    uint16_t *kind_state_branch = nullptr;
    float4 *kind_coef4 = nullptr, *kind_state4 = nullptr;  // float handles (see StepArgs::kind_coef4)
    uint32_t kinds = 0, kind_links = 0, kind_states = 0;
    size_t state_slots = 0;   // the states' arrays' length (general slabs, then face states)
    uint32_t face_specials = 0;
    uint32_t *face_tile = nullptr;
    uint4 *face_record = nullptr;
    uint16_t *face_state_kind = nullptr;
    float4 *face_coef = nullptr, *face_branch = nullptr;
    int face_lookup = 0;  // the face kinds' table has more than 255 rows
    int split = 0;            // 1: four launches a step instead of one
    int tile_order = 0;       // 0: tiles by die (see build), 1: in grid order, 2: by die, columns first
    int no_faces = 0;         // 1: every special is general (the adjoint's handles)
    int face_blocks = 0;      // 1: face specials by blocks of their own, 0: in their tiles
    uint32_t *face_tile_none = nullptr;  // a face_tile of empty ranges, for face_blocks
    // The forward run the adjoint goes back through (as_fdtd_record): the drive on the
    // device and per segment of rec_every steps the state before it (p(k - 1), p(k),
    // w(k - 1), w(k)).
    size_t rec_steps = 0, rec_every = 0;
    char *rec_ckpt = nullptr;
    float *rec_signal = nullptr;
    // With the history kept (keep_history), every step's p(n + 1) - p(n - 1) of each general
    // special (steps x specials, step-major), so that the adjoint needs no recompute; and
    // the record's graphs of rec_graph_every steps (by starting slot) on their staging.
    int keep_history = 0;  // 1: always, -1: never, 0: where it fits kHistoryBudget
    void *rec_hist = nullptr;
    hipGraphExec_t rec_graph[2] = {nullptr, nullptr};
    size_t rec_graph_every = 0;
    float *rec_sig_stage = nullptr;
    double *rec_out_stage = nullptr;
    void *rec_hist_stage = nullptr;
    std::vector<uint64_t> sorted_to_special;  // general special (sorted) -> layout index
    std::vector<uint64_t> state_slot;         // layout state -> its slot
    std::vector<uint32_t> probe_start_host;
    std::vector<uint32_t> probe_cell_host;
    std::vector<double> probe_weight_host;
    std::vector<double> coef_inv_host;        // per layout special
    std::vector<uint64_t> special_cell_host;  // padded cell per layout special
    // Where the grid has a matched layer, for its adjoint: its cells (padded) and each
    // layout special's axial link weights (0 where it has no link that way).
    std::vector<uint32_t> layer_cell_host;
    std::vector<float> special_axial_host;
    uint32_t *layer_cell = nullptr;
    uint8_t *layer_edge = nullptr;             // see StepArgs::layer_edge
    void *layer_face[2] = {nullptr, nullptr};  // the face velocities' slots; slot prev & 1 is current
    uint32_t px_magic = 0, px_shift = 0, ny_magic = 0, ny_shift = 0;
    float *layer_scale = nullptr;
    float *keep[6] = {}, *push[6] = {};
    uint32_t *probe_start = nullptr, *probe_cell = nullptr;
    double *probe_weight = nullptr;
    // staging for run() on host pointers
    float *signal_buf = nullptr;
    size_t signal_cap = 0;
    double *out_buf = nullptr;
    size_t out_cap = 0;
    // graphs of graph_steps steps (see ensure_graphs)
    uint32_t graph_steps = 64;
    hipGraphExec_t graph_exec[4] = {nullptr, nullptr, nullptr, nullptr};  // by the slot of p(k - 1) at start
    int graph_end[4] = {0, 0, 0, 0};                                        // and the slot after
    float *sig_stage = nullptr;
    double *out_stage = nullptr;
    std::string error;
};

namespace {

struct Failure {
    int code;
    std::string what;
};

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Failure{AS_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

// Captures what `enqueue` puts on `stream` into an executable graph at *exec. The capture is
// ended on every path: where `enqueue` throws, the partial graph is dropped and the stream
// leaves capture mode before the failure goes on, so the handle's stream stays usable.
template <typename F>
void capture_graph(hipStream_t stream, hipGraphExec_t *exec, F &&enqueue) {
    hipGraph_t graph = nullptr;
    check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        enqueue();
    } catch (...) {
        if (hipStreamEndCapture(stream, &graph) == hipSuccess && graph != nullptr) (void)hipGraphDestroy(graph);
        throw;
    }
    check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) *exec = nullptr;
    check(e, "hipGraphInstantiate");
}

template <typename U>
U *upload(as_fdtd *h, const std::vector<U> &v) {
    if (v.empty()) return nullptr;
    void *p = nullptr;
    check(hipMalloc(&p, v.size() * sizeof(U)), "hipMalloc");
    h->allocations.push_back(p);
    check(hipMemcpy(p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice), "hipMemcpy");
    return static_cast<U *>(p);
}

// A host array of doubles on the device as float (precision 32) or double (64).
void *upload_as(as_fdtd *h, const std::vector<double> &v, int precision) {
    if (precision == 64) return upload(h, v);
    std::vector<float> f(v.begin(), v.end());
    return upload(h, f);
}

void *zeros(as_fdtd *h, size_t bytes) {
    void *p = nullptr;
    check(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc");
    h->allocations.push_back(p);
    check(hipMemset(p, 0, std::max<size_t>(bytes, 256)), "hipMemset");
    return p;
}

void refuse(const std::string &what) { throw Failure{AS_EINVAL, what}; }

// work(slice, lo, hi) over [0, count) in slices on host threads (at most 16, and none for
// fewer than 2 * grain items), each slice once; the first exception thrown is rethrown.
// Returns the number of slices.
template <typename F>
unsigned in_slices(uint64_t count, uint64_t grain, F work) {
    const unsigned cores = std::max(1u, std::thread::hardware_concurrency());
    const unsigned slices = static_cast<unsigned>(std::min<uint64_t>({cores, 16u, std::max<uint64_t>(count / grain, 1)}));
    if (slices <= 1) {
        work(0u, uint64_t{0}, count);
        return 1;
    }
    std::vector<std::exception_ptr> errors(slices);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < slices; ++t) {
        pool.emplace_back([&, t] {
            try {
                work(t, count * t / slices, count * (t + 1) / slices);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        });
    }
    for (auto &th : pool) th.join();
    for (auto &e : errors)
        if (e) std::rethrow_exception(e);
    return slices;
}

// The set-up's phases' times on stderr, where AS_FDTD_TIMING is set.
struct PhaseClock {
    bool on = std::getenv("AS_FDTD_TIMING") != nullptr;
    std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    void mark(const char *what) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "as_fdtd_create: %-24s %9.1f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    }
};

// The multiplier and shift with which divide() divides by d (1 <= d < 2^31).
void divisor_of(uint32_t d, uint32_t *magic, uint32_t *shift) {
    uint32_t sh = 0;
    while ((1ull << sh) < d) ++sh;
    *shift = sh;
    *magic = static_cast<uint32_t>(((1ull << 32) * ((1ull << sh) - d)) / d + 1);
}

void build(as_fdtd *h, const as_layout &L, int precision, const as_tuning *tuning) {
    if (precision != 32 && precision != 64) refuse("precision is 32 or 64");
    if (L.stencil != 7 && L.stencil != 19) refuse("the stencil has 7 or 19 points");
    if (L.nx < 3 || L.ny < 3 || L.nz < 3) refuse("the grid is too small");
    if (L.n_weights < (L.stencil == 7 ? 6u : 18u)) refuse("the layout's weights are too few for its stencil");
    if (L.types == nullptr) refuse("the layout has no cell types");
    if (L.flow_mask != nullptr) {
        if (precision != 32) refuse("moving air is stepped in float only, the reference's scheme");
        const bool plain = L.flow_x != nullptr && L.flow_y != nullptr;
        const bool round = L.wind[0] != nullptr && L.wind[1] != nullptr && L.wind[2] != nullptr &&
                           L.loop[0] != nullptr && L.loop[1] != nullptr && L.loop[2] != nullptr;
        if (plain == round) refuse("moving air takes the shifts (flow_x, flow_y) or the wind and loops, not both");
    } else if (L.flow_x != nullptr || L.flow_y != nullptr || L.wind[0] != nullptr || L.loop[0] != nullptr) {
        refuse("moving air without its mask (flow_mask)");
    }
    // The probes' ranges: from 0, never decreasing, ending at the taps (a range running
    // backwards would make a probe's wavefront loop near 2^32 times; one past the taps would
    // read past the cells).
    if (L.probes > 0) {
        if (L.probe_start == nullptr) refuse("the layout's probes have no ranges (probe_start)");
        if (L.probe_start[0] != 0) refuse("the probes' ranges do not start at 0");
        for (uint64_t p = 0; p < L.probes; ++p)
            if (L.probe_start[p + 1] < L.probe_start[p]) refuse("the probes' ranges run backwards");
        if (L.probe_start[L.probes] != L.taps) refuse("the probes' ranges do not end at the taps");
    } else if (L.taps != 0) {
        refuse("the layout has taps but no probes");
    }
    if (L.taps > 0 && (L.probe_cells == nullptr || L.probe_weights == nullptr)) refuse("the probes' taps are missing");
    if (L.taps >= (1ull << 32)) refuse("more than 2^32 taps");
    h->precision = precision;
    h->wide = L.stencil == 19;
    h->scaled = L.scale != nullptr;
    h->nx = static_cast<int>(L.nx);
    h->ny = static_cast<int>(L.ny);
    h->nz = static_cast<int>(L.nz);
    int align = tuning && tuning->row_align > 0 ? tuning->row_align : 64;
    if (align != 1 && align != 16 && align != 32 && align != 64) refuse("the row alignment is 1, 16, 32 or 64 cells");
    // A small grid (fewer than 2048 tiles of 8 planes: too few blocks of four wavefronts to
    // keep the chip's 256 CUs' memory pipes full) takes tiles of 4 planes, each stepping its
    // own face specials; a larger one tiles of 8 beside blocks of face specials.
    const uint64_t tiles8 = ((L.nx + kTileX - 1) / kTileX) * ((L.ny + kTileRows - 3) / kTileRows) * ((L.nz + 5) / 8);
    const bool small = tiles8 < 2048;
    h->tile_z = tuning && tuning->tile_z > 0 ? tuning->tile_z : (small ? 4 : 8);
    h->face_blocks = tuning && tuning->face_blocks != 0 ? (tuning->face_blocks > 0 ? 1 : 0) : (small ? 0 : 1);
    if (tuning && tuning->tile_y > 0 && tuning->tile_y != kTileRows) refuse("tile_y is 4 (one wavefront a row)");
    if (tuning && tuning->graph_steps != 0) h->graph_steps = tuning->graph_steps < 0 ? 0u : static_cast<uint32_t>(tuning->graph_steps);
    if (h->graph_steps > 4096) refuse("at most 4096 steps a graph");
    if (tuning && tuning->split > 0) h->split = 1;
    if (tuning && tuning->tile_order > 0) h->tile_order = tuning->tile_order == 2 ? 2 : 1;
    if (tuning && tuning->no_faces > 0) h->no_faces = 1;
    if (tuning) h->keep_history = tuning->keep_history > 0 ? 1 : (tuning->keep_history < 0 ? -1 : 0);
    // Two steps a launch leave no room between the steps for moving air's carrying.
    const bool pairs_on = tuning && tuning->pairs > 0 && L.flow_mask == nullptr;
    const uint64_t nx = L.nx, ny = L.ny, nz = L.nz, n = nx * ny * nz;
    const uint64_t px = (nx + static_cast<uint64_t>(align) - 1) / static_cast<uint64_t>(align) * static_cast<uint64_t>(align);
    const uint64_t n_pad = px * ny * nz;
    if (n_pad >= (1ull << 31) || nx >= (1u << 21) || ny >= (1u << 21) || nz >= (1u << 21)) {
        throw Failure{AS_ERANGE, "the grid has too many cells for 32-bit indices"};
    }
    h->px = static_cast<int>(px);
    h->sxy = static_cast<long long>(px * ny);
    h->n_pad = n_pad;
    const auto pad = [&](uint64_t i) -> uint64_t {
        const uint64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        return x + px * (y + ny * z);
    };
    h->axial = L.weights[0];
    h->diagonal = h->wide ? L.weights[6] : 0.0;
    h->centre = L.centre;
    h->total = L.total;

    PhaseClock clock;
    // Cell classes, padded.
    std::vector<uint8_t> types(n_pad, 0);
    std::vector<float> scale;
    if (h->scaled) scale.assign(n_pad, 1.0f);
    // By rows, planes split among threads: row (y, z) of the layout to row (y, z) of pitch px.
    std::vector<uint64_t> marked(16, 0);
    in_slices(nz, 8, [&](unsigned t, uint64_t z0, uint64_t z1) {
        uint64_t count = 0;
        for (uint64_t z = z0; z < z1; ++z)
            for (uint64_t y = 0; y < ny; ++y) {
                const uint64_t from = nx * (y + ny * z), to = px * (y + ny * z);
                for (uint64_t x = 0; x < nx; ++x) {
                    const uint8_t c = L.types[from + x];
                    if (c > 3) refuse("a cell's type is not 0, 1, 2 or 3");
                    types[to + x] = c;
                    count += c == 2;
                }
                if (h->scaled) std::copy(L.scale + from, L.scale + from + nx, scale.begin() + static_cast<ptrdiff_t>(to));
            }
        marked[t] = count;
    });
    uint64_t marked_special = 0;
    for (uint64_t m : marked) marked_special += m;

    // Specials: checked, then each either a face special (its links axial, each weighing
    // the interior's axial weight, in the order of the directions; no speed per cell; at
    // most 15 states), stepped in its tile, or a general one, stepped by its own blocks.
    const uint64_t K = L.specials;
    if (K > 0 && (L.special == nullptr || L.link_start == nullptr || L.state_start == nullptr)) {
        refuse("the layout's special cells are incomplete");
    }
    if (K >= (1ull << 31) || L.links >= (1ull << 32) || L.states >= (1ull << 32)) {
        throw Failure{AS_ERANGE, "too many special cells, links or states for 32-bit indices"};
    }
    if (K > 0 && (L.link_start[0] != 0 || L.state_start[0] != 0 || L.link_start[K] != L.links ||
                  L.state_start[K] != L.states)) {
        refuse("the layout's link or state ranges do not match their arrays");
    }
    if (marked_special != K) refuse("the layout's special cells and its cells of type 2 differ");
    const auto links_of = [&](uint64_t k) { return L.link_start[k + 1] - L.link_start[k]; };
    const auto states_of = [&](uint64_t k) { return L.state_start[k + 1] - L.state_start[k]; };
    std::vector<uint8_t> face(K, 0);
    std::vector<uint8_t> seen(K ? n_pad : 0, 0);
    const float axial = L.weights[0];
    // Double coefficients (precision 64 only): every special then stepped by its own blocks.
    const bool exact = L.coef_b64 != nullptr || L.coef_inv64 != nullptr || L.branch_gain64 != nullptr;
    if (exact && (precision != 64 || L.coef_b64 == nullptr || L.coef_inv64 == nullptr ||
                  (L.branches > 0 && L.branch_gain64 == nullptr))) {
        refuse("double coefficients need precision 64 and all of coef_b64, coef_inv64 and branch_gain64");
    }
    if (exact) h->no_faces = 1;
    for (uint64_t k = 0; k < K; ++k) {
        const int64_t i = L.special[k];
        if (i < 0 || static_cast<uint64_t>(i) >= n) refuse("a special cell lies outside the grid");
        const uint64_t x = static_cast<uint64_t>(i) % nx, y = (static_cast<uint64_t>(i) / nx) % ny,
                       z = static_cast<uint64_t>(i) / (nx * ny);
        if (x == 0 || y == 0 || z == 0 || x + 1 == nx || y + 1 == ny || z + 1 == nz) {
            refuse("a special cell lies on the grid's edge");
        }
        if (L.types[i] != 2) refuse("a special cell is not of type 2");
        const uint64_t p = pad(static_cast<uint64_t>(i));
        if (seen[p]) refuse("a cell is special twice");
        seen[p] = 1;
        if (L.link_start[k + 1] < L.link_start[k] || L.state_start[k + 1] < L.state_start[k]) {
            refuse("the layout's link or state ranges run backwards");
        }
        if (links_of(k) > 18) refuse("a special cell has more than 18 links");
        if (states_of(k) > 255) refuse("a special cell has more than 255 branch states");
        for (uint64_t l = L.link_start[k]; l < L.link_start[k + 1]; ++l) {
            if (L.link_direction[l] >= 18) refuse("a link has no direction");
        }
        for (uint64_t j = L.state_start[k]; j < L.state_start[k + 1]; ++j) {
            if (L.state_branch[j] < 0 || static_cast<uint64_t>(L.state_branch[j]) >= L.branches) {
                refuse("a branch state has no branch");
            }
        }
        bool ok = !h->no_faces && !h->wide && !h->scaled && states_of(k) <= 15;
        int last = -1;
        for (uint64_t l = L.link_start[k]; ok && l < L.link_start[k + 1]; ++l) {
            const int d = L.link_direction[l];
            ok = d < 6 && d > last && L.link_weight[l] == axial;
            last = d;
        }
        face[k] = ok ? 1 : 0;
        if (ok) types[p] = kClassFace;
    }
    (void)seen;

    clock.mark("classes, specials");
    // The tiles: those holding interior cells or face specials.
    h->tiles_x = static_cast<uint32_t>((nx + kTileX - 1) / kTileX);
    h->tiles_y = static_cast<uint32_t>((ny - 2 + kTileRows - 1) / kTileRows);
    const uint32_t tiles_z = static_cast<uint32_t>((nz - 2 + h->tile_z - 1) / h->tile_z);
    if (static_cast<uint64_t>(h->tiles_x) * h->tiles_y * tiles_z >= kAllInterior) refuse("too many tiles");
    const auto tile_box = [&](uint32_t id, uint64_t *lo, uint64_t *hi) {
        const uint32_t tx = id % h->tiles_x, ty = (id / h->tiles_x) % h->tiles_y, tz = id / (h->tiles_x * h->tiles_y);
        lo[2] = 1 + static_cast<uint64_t>(tz) * h->tile_z;
        hi[2] = std::min<uint64_t>(lo[2] + h->tile_z, nz - 1);
        lo[1] = 1 + static_cast<uint64_t>(ty) * kTileRows;
        hi[1] = std::min<uint64_t>(lo[1] + kTileRows, ny - 1);
        lo[0] = static_cast<uint64_t>(tx) * kTileX;
        hi[0] = std::min<uint64_t>(lo[0] + kTileX, nx);
    };
    // Each tile's counts of interior cells and face specials (tiles split among threads),
    // then the list in tile order.
    const uint32_t tile_total = h->tiles_x * h->tiles_y * tiles_z;
    std::vector<uint32_t> tile_inner(tile_total, 0), tile_faces(tile_total, 0);
    in_slices(tile_total, 64, [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t id = a; id < b; ++id) {
            uint64_t lo[3], hi[3];
            uint32_t inner = 0, faces = 0;
            tile_box(static_cast<uint32_t>(id), lo, hi);
            for (uint64_t z = lo[2]; z < hi[2]; ++z)
                for (uint64_t y = lo[1]; y < hi[1]; ++y)
                    for (uint64_t x = lo[0]; x < hi[0]; ++x) {
                        const uint8_t c = types[x + px * (y + ny * z)];
                        inner += c == kClassInterior;
                        faces += c == kClassFace;
                    }
            tile_inner[id] = inner;
            tile_faces[id] = faces;
        }
    });
    std::vector<uint32_t> tiles;
    uint64_t interior = 0;
    for (uint32_t id = 0; id < tile_total; ++id) {
        uint64_t lo[3], hi[3];
        tile_box(id, lo, hi);
        if (tile_inner[id] + tile_faces[id]) {
            const bool all = tile_inner[id] == (hi[2] - lo[2]) * (hi[1] - lo[1]) * (hi[0] - lo[0]);
            tiles.push_back(id | (all ? kAllInterior : 0u));
        }
        interior += tile_inner[id];
    }
    h->interior_cells = interior;
    // Blocks go to the eight dies round robin by index (block b to die b mod 8), each die
    // with its own L2: so that a die's tiles are neighbours, block r takes tile
    // (r mod 8) * per + r / 8 of the list (x fastest, then y, then z): die d steps the d-th
    // slab of planes, and only the slabs' boundary planes are read by two dies.
    if (h->tile_order == 2) {
        // Columns first: z fastest, then x, then y, so a tile's neighbours above and below
        // run beside it on its die and their shared planes stay in that die's L2.
        const auto column_key = [&](uint32_t entry) {
            const uint32_t id = entry & ~kAllInterior;
            const uint32_t tx = id % h->tiles_x, ty = (id / h->tiles_x) % h->tiles_y, tz = id / (h->tiles_x * h->tiles_y);
            return (static_cast<uint64_t>(ty) * h->tiles_x + tx) * tiles_z + tz;
        };
        std::stable_sort(tiles.begin(), tiles.end(), [&](uint32_t a, uint32_t b) { return column_key(a) < column_key(b); });
    }
    if (h->tile_order != 1 && !tiles.empty()) {
        const size_t per = (tiles.size() + kXcds - 1) / kXcds;
        std::vector<uint32_t> ordered(per * kXcds, kNoTile);
        for (size_t r = 0; r < ordered.size(); ++r) {
            const size_t idx = (r % kXcds) * per + r / kXcds;
            if (idx < tiles.size()) ordered[r] = tiles[idx];
        }
        tiles.swap(ordered);
    }
    h->tile_count = static_cast<uint32_t>(tiles.size());
    // Each block's rows' lanes whose column holds an interior cell in the tile's planes.
    std::vector<uint64_t> tile_lanes(static_cast<size_t>(h->tile_count) * kTileRows, 0);
    in_slices(h->tile_count, 64, [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t r = a; r < b; ++r) {
            if (tiles[r] == kNoTile) continue;
            uint64_t lo[3], hi[3];
            tile_box(tiles[r] & ~kAllInterior, lo, hi);
            for (uint64_t y = lo[1]; y < hi[1]; ++y) {
                uint64_t lanes = 0;
                for (uint64_t x = lo[0]; x < hi[0]; ++x)
                    for (uint64_t z = lo[2]; z < hi[2]; ++z)
                        if (types[x + px * (y + ny * z)] == kClassInterior) {
                            lanes |= uint64_t{1} << (x - lo[0]);
                            break;
                        }
                tile_lanes[r * kTileRows + (y - lo[1])] = lanes;
            }
        }
    });

    // Each row's interior cells as bits (see StepArgs::interior_bits), planes among threads.
    std::vector<uint64_t> interior_bits(static_cast<size_t>(nz) * ny * h->tiles_x, 0);
    in_slices(nz, 1, [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t z = a; z < b; ++z)
            for (uint64_t y = 0; y < ny; ++y)
                for (uint64_t x = 0; x < nx; ++x)
                    if (types[x + px * (y + ny * z)] == kClassInterior)
                        interior_bits[(z * ny + y) * h->tiles_x + x / kTileX] |= uint64_t{1} << (x % kTileX);
    });

    // General specials: in the order of the interior tiles holding them, or by their counts
    // of links and states, or by cell (a special's update reads only its own cell, links and
    // states, so their order is free), in blocks of kBlock;
    // each block's links and states in slabs of (its largest count) x kBlock, slot-major.
    // A block's links are its chunks' packed directions and the links' weights, slot-major;
    // each state names its branch, whose coefficients sit in small tables.
    clock.mark("tiles");
    if (L.branches > 0xFFFFu) refuse("more than 65535 branches");
    std::vector<uint64_t> order;
    for (uint64_t k = 0; k < K; ++k)
        if (!face[k]) order.push_back(k);
    const uint64_t G = order.size();
    // By tile by default: a block's specials then lie close together, and their gathers
    // share cache lines (the fitted hall: 34.96 to 32.10 us a step against by counts).
    const int special_order = tuning && tuning->special_order > 0 ? tuning->special_order : 3;
    if (special_order == 1) {
        std::stable_sort(order.begin(), order.end(), [&](uint64_t p, uint64_t q) {
            return links_of(p) != links_of(q) ? links_of(p) < links_of(q) : states_of(p) < states_of(q);
        });
    } else if (special_order == 3) {
        // By the interior tile (64 x 4 x 8 cells) that holds the cell, tiles in grid order.
        const auto tile_key = [&](uint64_t k) {
            const uint64_t c = static_cast<uint64_t>(L.special[k]);
            const uint64_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
            return ((z / 8) * ((ny + 3) / 4) + y / 4) * ((nx + 63) / 64) + x / 64;
        };
        std::stable_sort(order.begin(), order.end(), [&](uint64_t p, uint64_t q) { return tile_key(p) < tile_key(q); });
    }
    // special_order 2: the layout's order (ascending cells, then the source's).
    const uint64_t blocks = (G + kBlock - 1) / kBlock;
    std::vector<uint32_t> sp_cell(G), link_base(blocks), state_base(blocks);
    std::vector<uint8_t> sp_links(G), sp_states(G);
    std::vector<float> coef_a(G), source(G);
    std::vector<double> coef_b(G), coef_inv(G);
    uint64_t link_slots = 0, state_slots = 0;
    for (uint64_t B = 0; B < blocks; ++B) {
        uint64_t wl = 0, ws = 0;
        for (uint64_t t = 0; t < kBlock && B * kBlock + t < G; ++t) {
            wl = std::max(wl, links_of(order[B * kBlock + t]));
            ws = std::max(ws, states_of(order[B * kBlock + t]));
        }
        wl = (wl + kLinkChunk - 1) / kLinkChunk * kLinkChunk;  // whole chunks: link_base / kLinkChunk
        link_base[B] = static_cast<uint32_t>(link_slots);
        state_base[B] = static_cast<uint32_t>(state_slots);
        link_slots += wl * kBlock;
        state_slots += ws * kBlock;
        if (link_slots >= (1ull << 32) || state_slots >= (1ull << 31)) {
            throw Failure{AS_ERANGE, "too many links or states for 32-bit indices"};
        }
    }
    std::vector<uint32_t> link_dirs(link_slots / kLinkChunk, 0);
    std::vector<float> link_weight(link_slots, 0.0f);
    std::vector<float> state_weight(state_slots, 0.0f);
    std::vector<uint16_t> state_branch(state_slots, 0);
    for (uint64_t k = 0; k < G; ++k) {
        const uint64_t src = order[k], B = k / kBlock, t = k % kBlock;
        sp_cell[k] = static_cast<uint32_t>(pad(static_cast<uint64_t>(L.special[src])));
        sp_links[k] = static_cast<uint8_t>(links_of(src));
        sp_states[k] = static_cast<uint8_t>(states_of(src));
        coef_a[k] = L.coef_a[src];
        coef_b[k] = exact ? L.coef_b64[src] : L.coef_b[src];
        coef_inv[k] = exact ? L.coef_inv64[src] : L.coef_inv[src];
        source[k] = L.source[src];
        for (uint64_t u = 0; u < links_of(src); ++u) {
            const uint64_t l = L.link_start[src] + u;
            const uint64_t slot = link_base[B] + u * kBlock + t;
            link_dirs[link_base[B] / kLinkChunk + (u / kLinkChunk) * kBlock + t] |=
                static_cast<uint32_t>(L.link_direction[l]) << (5 * (u % kLinkChunk));
            link_weight[slot] = L.link_weight[l];
        }
        for (uint64_t u = 0; u < states_of(src); ++u) {
            const uint64_t j = L.state_start[src] + u;
            const uint64_t slot = state_base[B] + u * kBlock + t;
            state_weight[slot] = L.state_weight[j];
            state_branch[slot] = static_cast<uint16_t>(L.state_branch[j]);
        }
    }
    // The specials' kinds: specials alike in everything but their cell and their states'
    // velocities (coefficients, links' directions and weights, states' branches and
    // weights, bit for bit) share one row of the kinds' tables. Used where a kind serves
    // four specials or more on average (the cut cells of a plane repeat along the plane's
    // axes: the fitted hall's 486,785 general specials are 5,003 kinds), or as asked.
    std::vector<uint32_t> sp_kind(G);
    std::vector<uint64_t> kind_first;  // per kind: a sorted special of it
    {
        // Each special's fields hashed on the set-up's threads; then, in order, a special
        // joins the first kind of equal hash whose fields equal its own (bit for bit), or
        // starts a kind.
        const auto bits32 = [](float v) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            return u;
        };
        const auto bits64 = [](double v) {
            uint64_t u;
            std::memcpy(&u, &v, 8);
            return u;
        };
        const auto same = [&](uint64_t p, uint64_t q) {
            const uint64_t a = order[p], c = order[q];
            if (bits32(coef_a[p]) != bits32(coef_a[q]) || bits64(coef_b[p]) != bits64(coef_b[q]) ||
                bits64(coef_inv[p]) != bits64(coef_inv[q]) || bits32(source[p]) != bits32(source[q]) ||
                links_of(a) != links_of(c) || states_of(a) != states_of(c))
                return false;
            for (uint64_t u = 0; u < links_of(a); ++u) {
                const uint64_t l = L.link_start[a] + u, m = L.link_start[c] + u;
                if (L.link_direction[l] != L.link_direction[m] || bits32(L.link_weight[l]) != bits32(L.link_weight[m]))
                    return false;
            }
            for (uint64_t u = 0; u < states_of(a); ++u) {
                const uint64_t j = L.state_start[a] + u, m = L.state_start[c] + u;
                if (L.state_branch[j] != L.state_branch[m] || bits32(L.state_weight[j]) != bits32(L.state_weight[m]))
                    return false;
            }
            return true;
        };
        std::vector<uint64_t> hash(G);
        in_slices(G, 4096, [&](unsigned, uint64_t a, uint64_t b) {
            for (uint64_t k = a; k < b; ++k) {
                const uint64_t src = order[k];
                uint64_t h = 0x9E3779B97F4A7C15ull;
                const auto mix = [&h](uint64_t v) {
                    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
                    h *= 0xBF58476D1CE4E5B9ull;
                    h ^= h >> 31;
                };
                mix(bits32(coef_a[k]));
                mix(bits64(coef_b[k]));
                mix(bits64(coef_inv[k]));
                mix(bits32(source[k]));
                mix(links_of(src) | states_of(src) << 32);
                for (uint64_t u = 0; u < links_of(src); ++u) {
                    const uint64_t l = L.link_start[src] + u;
                    mix(static_cast<uint64_t>(L.link_direction[l]) << 32 | bits32(L.link_weight[l]));
                }
                for (uint64_t u = 0; u < states_of(src); ++u) {
                    const uint64_t j = L.state_start[src] + u;
                    mix(static_cast<uint64_t>(static_cast<uint32_t>(L.state_branch[j])) << 32 | bits32(L.state_weight[j]));
                }
                hash[k] = h;
            }
        });
        std::unordered_map<uint64_t, uint32_t> first_of;  // hash -> its first kind
        std::vector<uint32_t> also;                       // per kind: the next kind of its hash
        for (uint64_t k = 0; k < G; ++k) {
            const auto at = first_of.find(hash[k]);
            uint32_t c = at == first_of.end() ? ~0u : at->second, last = ~0u;
            while (c != ~0u && !same(kind_first[c], k)) {
                last = c;
                c = also[c];
            }
            if (c == ~0u) {
                c = static_cast<uint32_t>(kind_first.size());
                kind_first.push_back(k);
                also.push_back(~0u);
                if (last == ~0u)
                    first_of.emplace(hash[k], c);
                else
                    also[last] = c;
            }
            sp_kind[k] = c;
        }
    }
    const int kinds_asked = tuning ? tuning->special_kinds : 0;
    const bool use_kinds = G > 0 && (kinds_asked == 2 || (kinds_asked == 0 && kind_first.size() * 4 <= G));
    std::vector<uint4> sp_rec;
    std::vector<float> kind_coef_a, kind_source, kind_weight, kind_state_weight;
    std::vector<double> kind_coef_b, kind_coef_inv;
    std::vector<uint32_t> kind_dirs;
    std::vector<uint16_t> kind_state_branch;
    std::vector<float4> kind_coef4, kind_state4;  // float handles: (coef_a, coef_b, coef_inv, source), (now, before, gain, weight)
    if (use_kinds) {
        const uint64_t nk = kind_first.size();
        uint64_t wl = 0, ws = 0;
        for (uint64_t q : kind_first) {
            wl = std::max(wl, links_of(order[q]));
            ws = std::max(ws, states_of(order[q]));
        }
        // Whole chunks of directions, and whole float4 rows of weights (see special_one).
        wl = std::max<uint64_t>((wl + 2 * kLinkChunk - 1) / (2 * kLinkChunk) * (2 * kLinkChunk), 2 * kLinkChunk);
        ws = std::max<uint64_t>(ws, 1);
        if (nk * wl >= (1ull << 32) || nk * ws >= (1ull << 32))
            throw Failure{AS_ERANGE, "too many links or states in the specials' kinds for 32-bit indices"};
        kind_coef_a.resize(nk);
        kind_source.resize(nk);
        kind_coef_b.resize(nk);
        kind_coef_inv.resize(nk);
        kind_weight.assign(nk * wl, 0.0f);
        kind_dirs.assign(nk * wl / kLinkChunk, 0u);
        kind_state_weight.assign(nk * ws, 0.0f);
        kind_state_branch.assign(nk * ws, 0);
        kind_coef4.assign(nk, make_float4(0, 0, 0, 0));
        kind_state4.assign(nk * ws, make_float4(0, 0, 0, 0));
        for (uint64_t c = 0; c < nk; ++c) {
            const uint64_t k = kind_first[c], src = order[k];
            kind_coef_a[c] = coef_a[k];
            kind_source[c] = source[k];
            kind_coef_b[c] = coef_b[k];
            kind_coef_inv[c] = coef_inv[k];
            for (uint64_t u = 0; u < links_of(src); ++u) {
                const uint64_t l = L.link_start[src] + u;
                kind_weight[c * wl + u] = L.link_weight[l];
                kind_dirs[c * wl / kLinkChunk + u / kLinkChunk] |= static_cast<uint32_t>(L.link_direction[l])
                                                                   << (5 * (u % kLinkChunk));
            }
            for (uint64_t u = 0; u < states_of(src); ++u) {
                const uint64_t j = L.state_start[src] + u;
                kind_state_weight[c * ws + u] = L.state_weight[j];
                kind_state_branch[c * ws + u] = static_cast<uint16_t>(L.state_branch[j]);
                const int32_t br = L.state_branch[j];
                kind_state4[c * ws + u] =
                    make_float4(L.branch_now[br], L.branch_before[br], L.branch_gain[br], L.state_weight[j]);
            }
            kind_coef4[c] = make_float4(coef_a[k], static_cast<float>(coef_b[k]), static_cast<float>(coef_inv[k]),
                                        source[k]);
        }
        sp_rec.resize(G);
        for (uint64_t k = 0; k < G; ++k) {
            const uint32_t c = sp_kind[k];
            sp_rec[k] = make_uint4(sp_cell[k], c, kind_dirs[c * wl / kLinkChunk],
                                   static_cast<uint32_t>(sp_links[k]) | static_cast<uint32_t>(sp_states[k]) << 8);
        }
        h->kinds = static_cast<uint32_t>(nk);
        h->kind_links = static_cast<uint32_t>(wl);
        h->kind_states = static_cast<uint32_t>(ws);
    }
    h->specials = static_cast<uint32_t>(G);
    // Where each layout special and state went, for the adjoint's sensitivities.
    h->sorted_to_special = order;
    h->state_slot.assign(L.states, ~0ull);
    h->coef_inv_host.resize(K);
    for (uint64_t k = 0; k < K; ++k) h->coef_inv_host[k] = exact ? L.coef_inv64[k] : L.coef_inv[k];
    h->special_cell_host.resize(K);
    for (uint64_t k = 0; k < K; ++k) h->special_cell_host[k] = pad(static_cast<uint64_t>(L.special[k]));
    if (L.layer_cells > 0) {
        h->special_axial_host.assign(6 * K, 0.0f);
        for (uint64_t k = 0; k < K; ++k)
            for (uint64_t l = L.link_start[k]; l < L.link_start[k + 1]; ++l)
                if (L.link_direction[l] < 6) h->special_axial_host[6 * k + L.link_direction[l]] = L.link_weight[l];
    }
    for (uint64_t k = 0; k < G; ++k) {
        const uint64_t src = order[k], B = k / kBlock, t = k % kBlock;
        for (uint64_t u = 0; u < states_of(src); ++u) h->state_slot[L.state_start[src] + u] = state_base[B] + u * kBlock + t;
    }
    h->face_specials = static_cast<uint32_t>(K - G);
    h->states = static_cast<uint32_t>(L.states);

    // Face specials' records, tile block by tile block (each tile's in the order of its
    // cells). Their states follow the general specials' slabs in the states' arrays.
    std::vector<std::pair<uint64_t, uint64_t>> by_cell;  // (padded cell, special)
    for (uint64_t k = 0; k < K; ++k)
        if (face[k]) by_cell.emplace_back(pad(static_cast<uint64_t>(L.special[k])), k);
    std::sort(by_cell.begin(), by_cell.end());
    const auto bits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    // A row's index in its table (the rows numbered in the order first met), through a small
    // direct-mapped cache of recent rows: walls repeat their coefficients and kinds.
    struct RowTable {
        std::map<std::array<uint32_t, 4>, uint32_t> index;
        std::array<std::pair<std::array<uint32_t, 4>, uint32_t>, 64> recent{};
        std::array<bool, 64> held{};
    };
    RowTable coef_index, branch_index;
    std::vector<float> face_coef, face_branch;
    const auto index_of = [&](RowTable &table, std::vector<float> &rows, const float (&v)[4]) {
        const std::array<uint32_t, 4> key{bits(v[0]), bits(v[1]), bits(v[2]), bits(v[3])};
        const uint32_t slot = (key[0] ^ (key[1] * 0x9E3779B1u) ^ (key[2] * 0x85EBCA77u) ^ (key[3] * 0xC2B2AE3Du)) >> 26;
        if (table.held[slot] && table.recent[slot].first == key) return table.recent[slot].second;
        uint32_t id;
        const auto it = table.index.find(key);
        if (it != table.index.end()) {
            id = it->second;
        } else {
            id = static_cast<uint32_t>(table.index.size());
            table.index.emplace(key, id);
            rows.insert(rows.end(), v, v + 4);
        }
        table.recent[slot] = {key, id};
        table.held[slot] = true;
        return id;
    };
    std::vector<uint32_t> face_tile(static_cast<size_t>(h->tile_count) + 1, 0), face_cell, face_desc, face_state, face_kinds;
    std::vector<uint16_t> face_state_kind;
    std::vector<size_t> face_of(K, ~size_t{0});  // special -> its record
    uint64_t face_slots = state_slots;
    // Each face special's tile block, then the face specials by block (a counting sort that
    // keeps each block's in ascending cell order, the order of its cells in the tile).
    std::vector<uint32_t> block_of_tile(tile_total, kNoTile);
    for (uint32_t blk = 0; blk < h->tile_count; ++blk) {
        if (tiles[blk] != kNoTile) block_of_tile[tiles[blk] & ~kAllInterior] = blk;
    }
    std::vector<uint32_t> face_block(by_cell.size());
    for (size_t f = 0; f < by_cell.size(); ++f) {
        const uint64_t p = by_cell[f].first, x = p % px, y = (p / px) % ny, z = p / (px * ny);
        const uint64_t id = x / kTileX + h->tiles_x * ((y - 1) / kTileRows + h->tiles_y * ((z - 1) / static_cast<uint64_t>(h->tile_z)));
        face_block[f] = block_of_tile[id];
        if (face_block[f] == kNoTile) throw Failure{AS_EINVAL, "internal: a face special was not met by a tile"};
        ++face_tile[face_block[f] + 1];
    }
    for (uint32_t blk = 0; blk < h->tile_count; ++blk) face_tile[blk + 1] += face_tile[blk];
    std::vector<uint64_t> in_block_order(by_cell.size());
    {
        std::vector<uint32_t> next(face_tile.begin(), face_tile.end() - 1);
        for (size_t f = 0; f < by_cell.size(); ++f) in_block_order[next[face_block[f]]++] = f;
    }
    for (const uint64_t f : in_block_order) {
        const uint64_t p = by_cell[f].first, k = by_cell[f].second;
        uint32_t mask = 0;
        for (uint64_t l = L.link_start[k]; l < L.link_start[k + 1]; ++l) mask |= 1u << L.link_direction[l];
        const float cv[4] = {L.coef_a[k], L.coef_b[k], L.coef_inv[k], L.source[k]};
        const uint32_t ci = index_of(coef_index, face_coef, cv);
        if (ci >= (1u << 22)) throw Failure{AS_ERANGE, "too many distinct special coefficients"};
        face_cell.push_back(static_cast<uint32_t>(p));
        face_desc.push_back(mask | static_cast<uint32_t>(states_of(k)) << 6 | ci << 10);
        face_state.push_back(static_cast<uint32_t>(face_slots));
        face_of[k] = face_desc.size() - 1;
        uint32_t packed = 0;
        for (uint64_t j = L.state_start[k]; j < L.state_start[k + 1]; ++j) {
            const int32_t br = L.state_branch[j];
            const float bv[4] = {L.branch_now[br], L.branch_before[br], L.branch_gain[br], L.state_weight[j]};
            const uint32_t bi = index_of(branch_index, face_branch, bv);
            if (bi > 0xFFFFu) throw Failure{AS_ERANGE, "too many distinct branch states"};
            const uint64_t u = j - L.state_start[k];
            if (u < 4) packed |= std::min<uint32_t>(bi, 0xFFu) << (8 * u);
            face_state_kind.push_back(static_cast<uint16_t>(bi));
            ++face_slots;
        }
        for (uint64_t u = states_of(k); u < 4; ++u) packed |= 0xFFu << (8 * u);
        face_kinds.push_back(packed);
        if (face_slots >= (1ull << 31)) throw Failure{AS_ERANGE, "too many states"};
    }
    if (face_desc.size() != K - G) throw Failure{AS_EINVAL, "internal: a face special was not met by a tile"};
    // The face states' kinds are indexed from the start of the states' arrays.
    std::vector<uint16_t> kinds(face_slots, 0);
    std::copy(face_state_kind.begin(), face_state_kind.end(), kinds.begin() + static_cast<ptrdiff_t>(state_slots));
    h->state_slots = face_slots;

    clock.mark("slabs, face records");
    // Two steps a launch (pair_kernel): float, 7 points, one speed of sound, and arrays
    // within a buffer resource's 32-bit byte range.
    const uint64_t groups = (ny - 2 + kBoxRows - 1) / kBoxRows;
    h->pairs = pairs_on && precision == 32 && !h->wide && !h->scaled && nz >= 10 && n_pad * sizeof(float) < (1ull << 32) &&
                       nz * groups * px * sizeof(uint32_t) < (1ull << 32)
                   ? 1
                   : 0;
    std::vector<uint32_t> boxes, box_bits, cells;
    if (h->pairs) {
        const uint64_t boxes_x = (nx + kBoxX - 1) / kBoxX;
        const auto count_boxes = [&](uint64_t Z) { return boxes_x * groups * ((nz - 2 + Z - 1) / Z); };
        // Boxes of 8 planes (on the hall, 16 planes' boxes are slower: fewer wavefronts to
        // hide a plane's loads), or 16 as asked (pairs 16) where the grid has the planes.
        int Z = 8;
        if ((tuning->pairs == 8 || tuning->pairs == 16) && nz - 2 >= static_cast<uint64_t>(tuning->pairs)) Z = tuning->pairs;
        const uint64_t boxes_z = (nz - 2 + Z - 1) / Z;
        if (count_boxes(Z) >= kNoTile) refuse("too many boxes");
        h->pair_z = Z;
        h->boxes_x = static_cast<uint32_t>(boxes_x);
        h->groups_y = static_cast<uint32_t>(groups);
        // Each interior cell's bits (interior; deep: its six neighbours interior too); the
        // interior cells that are not deep go to the list the second launch steps.
        box_bits.assign(nz * groups * px, 0);
        std::vector<uint8_t> active(count_boxes(Z), 0);
        const auto inner = [&](uint64_t x, uint64_t y, uint64_t z) {
            return types[x + px * (y + ny * z)] == kClassInterior;
        };
        bool edge = false;  // an interior cell on the grid's edge: one step a launch only
        for (uint64_t z = 0; z < nz && !edge; ++z)
            for (uint64_t y = 0; y < ny && !edge; ++y)
                for (uint64_t x = 0; x < nx; ++x) {
                    if (!inner(x, y, z)) continue;
                    if (x == 0 || y == 0 || z == 0 || x + 1 == nx || y + 1 == ny || z + 1 == nz) {
                        edge = true;
                        break;
                    }
                    const bool deep = inner(x - 1, y, z) && inner(x + 1, y, z) && inner(x, y - 1, z) &&
                                      inner(x, y + 1, z) && inner(x, y, z - 1) && inner(x, y, z + 1);
                    const uint64_t g = (y - 1) / kBoxRows, r = (y - 1) % kBoxRows;
                    box_bits[(z * groups + g) * px + x] |= (deep ? 3u : 1u) << (8 * r);
                    if (!deep) cells.push_back(static_cast<uint32_t>(x + px * (y + ny * z)));
                    const uint64_t bz = std::min<uint64_t>((z - 1) / static_cast<uint64_t>(Z), boxes_z - 1);
                    active[g + groups * (x / kBoxX + boxes_x * bz)] = 1;
                }
        if (edge) {
            h->pairs = 0;
            box_bits.clear();
            cells.clear();
        } else {
            // The boxes in order (row groups fastest, then x, then z), kBoxWaves a block; block
            // r takes chunk (r mod 8) * per + r / 8, so that each die's blocks are neighbours
            // (see the tiles).
            std::vector<uint32_t> list;
            for (uint64_t b = 0; b < active.size(); ++b)
                if (active[b]) list.push_back(static_cast<uint32_t>(b));
            const size_t chunks = (list.size() + kBoxWaves - 1) / kBoxWaves;
            const size_t per = (chunks + kXcds - 1) / kXcds;
            boxes.assign(per * kXcds * kBoxWaves, kNoTile);
            for (size_t r = 0; r < per * kXcds; ++r) {
                const size_t chunk = h->tile_order != 1 ? (r % kXcds) * per + r / kXcds : r;
                for (uint32_t w = 0; w < kBoxWaves; ++w) {
                    const size_t at = chunk * kBoxWaves + w;
                    if (at < list.size()) boxes[r * kBoxWaves + w] = list[at];
                }
            }
            h->box_blocks = static_cast<uint32_t>(per * kXcds);
            h->cell_count = static_cast<uint32_t>(cells.size());
        }
    }

    clock.mark("boxes");
    // Matched layer.
    std::vector<uint32_t> layer_cell(L.layer_cells);
    in_slices(L.layer_cells, 1 << 16, [&](unsigned, uint64_t a, uint64_t b) {
        for (uint64_t c = a; c < b; ++c) {
            const int64_t i = L.layer_cell[c];
            if (i < 0 || static_cast<uint64_t>(i) >= n) refuse("a layer cell lies outside the grid");
            const uint64_t x = static_cast<uint64_t>(i) % nx, y = (static_cast<uint64_t>(i) / nx) % ny,
                           z = static_cast<uint64_t>(i) / (nx * ny);
            if (x == 0 || y == 0 || z == 0 || x + 1 == nx || y + 1 == ny || z + 1 == nz) {
                refuse("a layer cell lies on the grid's edge");
            }
            layer_cell[c] = static_cast<uint32_t>(x + px * (y + ny * z));
        }
    });
    h->layer_cells = static_cast<uint32_t>(L.layer_cells);
    // Which neighbours below a layer cell are layer cells too (their shared face is theirs).
    std::vector<uint8_t> layer_edge(L.layer_cells);
    if (L.layer_cells > 0) {
        std::vector<uint8_t> member(n_pad, 0);
        in_slices(L.layer_cells, 1 << 16, [&](unsigned, uint64_t a, uint64_t b) {
            for (uint64_t c = a; c < b; ++c) member[layer_cell[c]] = 1;
        });
        const uint64_t below[3] = {1, px, px * ny};
        in_slices(L.layer_cells, 1 << 16, [&](unsigned, uint64_t a, uint64_t b) {
            for (uint64_t c = a; c < b; ++c) {
                uint8_t bits = 0;
                for (int axis = 0; axis < 3; ++axis) {
                    if (!member[layer_cell[c] - below[axis]]) bits |= static_cast<uint8_t>(1u << axis);
                }
                layer_edge[c] = bits;
            }
        });
    }

    // Probes.
    std::vector<uint32_t> probe_start(L.probes + 1), probe_cell(L.taps);
    std::vector<double> probe_weight(L.probe_weights, L.probe_weights + L.taps);
    for (uint64_t p = 0; p <= L.probes; ++p) probe_start[p] = static_cast<uint32_t>(L.probe_start[p]);
    for (uint64_t t = 0; t < L.taps; ++t) {
        if (L.probe_cells[t] >= n) refuse("a probe reads a cell outside the grid");
        probe_cell[t] = static_cast<uint32_t>(pad(L.probe_cells[t]));
    }
    h->probes = static_cast<uint32_t>(L.probes);
    h->probe_start_host = probe_start;
    h->probe_cell_host = probe_cell;
    h->probe_weight_host = probe_weight;

    clock.mark("layer, probes");
    // Device copies.
    const size_t e = precision == 64 ? sizeof(double) : sizeof(float);
    check(hipGetDevice(&h->device), "hipGetDevice");
    check(hipStreamCreate(&h->stream), "hipStreamCreate");
    h->types = upload(h, types);
    h->scale = upload(h, scale);
    h->tiles = upload(h, tiles);
    h->tile_lanes = upload(h, tile_lanes);
    h->interior_bits = upload(h, interior_bits);
    h->sp_cell = upload(h, sp_cell);
    h->sp_links = upload(h, sp_links);
    h->sp_states = upload(h, sp_states);
    h->link_base = upload(h, link_base);
    h->state_base = upload(h, state_base);
    h->link_dirs = upload(h, link_dirs);
    h->link_weight = upload(h, link_weight);
    h->coef_a = upload(h, coef_a);
    h->coef_b = upload_as(h, coef_b, precision);
    h->coef_inv = upload_as(h, coef_inv, precision);
    h->source = upload(h, source);
    h->state_weight = upload(h, state_weight);
    h->state_branch = upload(h, state_branch);
    h->sp_rec = upload(h, sp_rec);
    h->kind_coef_a = upload(h, kind_coef_a);
    h->kind_source = upload(h, kind_source);
    h->kind_coef_b = kind_coef_b.empty() ? nullptr : upload_as(h, kind_coef_b, precision);
    h->kind_coef_inv = kind_coef_inv.empty() ? nullptr : upload_as(h, kind_coef_inv, precision);
    h->kind_weight = upload(h, kind_weight);
    h->kind_dirs = upload(h, kind_dirs);
    h->kind_state_weight = upload(h, kind_state_weight);
    h->kind_state_branch = upload(h, kind_state_branch);
    if (precision == 32) {
        h->kind_coef4 = upload(h, kind_coef4);
        h->kind_state4 = upload(h, kind_state4);
    }
    {
        std::vector<float> now(L.branch_now, L.branch_now + L.branches), before(L.branch_before, L.branch_before + L.branches);
        std::vector<double> gain(L.branches);
        for (uint64_t br = 0; br < L.branches; ++br) gain[br] = exact ? L.branch_gain64[br] : L.branch_gain[br];
        h->branch_now = upload(h, now);
        h->branch_before = upload(h, before);
        h->branch_gain = upload_as(h, gain, precision);
    }
    h->face_tile = upload(h, face_tile);
    h->face_tile_none = upload(h, std::vector<uint32_t>(face_tile.size(), 0));
    {
        std::vector<uint32_t> records(4 * face_cell.size());
        for (size_t r = 0; r < face_cell.size(); ++r) {
            records[4 * r] = face_cell[r];
            records[4 * r + 1] = face_desc[r];
            records[4 * r + 2] = face_state[r];
            records[4 * r + 3] = face_kinds[r];
        }
        h->face_record = reinterpret_cast<uint4 *>(upload(h, records));
    }
    h->face_state_kind = upload(h, kinds);
    h->face_coef = reinterpret_cast<float4 *>(upload(h, face_coef));
    // Kinds that all fit a byte read row 255 for the state slots a special lacks: pad the
    // table to 256 rows of which the last is zeros. A larger table keeps the look-ups.
    h->face_lookup = face_branch.size() / 4 >= 255 ? 1 : 0;
    if (!h->face_lookup && !face_desc.empty()) face_branch.resize(256 * 4, 0.0f);
    h->face_branch = reinterpret_cast<float4 *>(upload(h, face_branch));
    h->layer_cell = upload(h, layer_cell);
    h->layer_cell_host = layer_cell;
    h->layer_edge = upload(h, layer_edge);
    divisor_of(static_cast<uint32_t>(px), &h->px_magic, &h->px_shift);
    divisor_of(static_cast<uint32_t>(ny), &h->ny_magic, &h->ny_shift);
    if (L.layer_scale != nullptr && L.layer_cells > 0) {
        h->layer_scale = upload(h, std::vector<float>(L.layer_scale, L.layer_scale + L.layer_cells));
    }
    if (L.layer_cells > 0) {
        const uint64_t len[3] = {nx + 1, ny + 1, nz + 1};
        for (int k = 0; k < 6; ++k) {
            if (L.keep[k] == nullptr || L.push[k] == nullptr) refuse("the matched layer's coefficients are missing");
            h->keep[k] = upload(h, std::vector<float>(L.keep[k], L.keep[k] + len[k / 2]));
            h->push[k] = upload(h, std::vector<float>(L.push[k], L.push[k] + len[k / 2]));
        }
    }
    h->probe_start = upload(h, probe_start);
    h->probe_cell = upload(h, probe_cell);
    h->probe_weight = upload(h, probe_weight);
    h->u0 = zeros(h, n_pad * e);
    h->u1 = zeros(h, n_pad * e);
    h->w0 = zeros(h, face_slots * e);
    h->w1 = zeros(h, face_slots * e);
    h->field[0] = h->u0;
    h->field[1] = h->u1;
    h->velocity[0] = h->w0;
    h->velocity[1] = h->w1;
    if (h->pairs) {
        for (int k = 2; k < 4; ++k) {
            h->field[k] = zeros(h, n_pad * e);
            h->velocity[k] = zeros(h, face_slots * e);
        }
        h->boxes = upload(h, boxes);
        h->box_bits = upload(h, box_bits);
        h->cells = upload(h, cells);
    }
    if (L.flow_mask != nullptr) {
        // Moving air, padded: cells off the grid's rows read 0 (solid).
        const auto padded = [&](const auto *src) {
            using U = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
            std::vector<U> out(n_pad, U{});
            for (uint64_t z = 0; z < nz; ++z)
                for (uint64_t y = 0; y < ny; ++y)
                    std::copy(src + nx * (y + ny * z), src + nx * (y + ny * z) + nx,
                              out.begin() + static_cast<ptrdiff_t>(px * (y + ny * z)));
            return out;
        };
        h->flowing = true;
        h->flow_round = L.wind[0] != nullptr;
        h->flow_mask = upload(h, padded(L.flow_mask));
        if (h->flow_round) {
            for (int d = 0; d < 3; ++d) {
                h->flow_wind[d] = upload(h, padded(L.wind[d]));
                h->flow_loop[d] = upload(h, padded(L.loop[d]));
            }
        } else {
            h->flow_x = upload(h, padded(L.flow_x));
            h->flow_y = upload(h, padded(L.flow_y));
        }
        h->flow_t1 = static_cast<float *>(zeros(h, n_pad * sizeof(float)));
        h->flow_t2 = static_cast<float *>(zeros(h, n_pad * sizeof(float)));
        // A slot a block of the carrying's launches: a cell a thread, or a plain tile a block.
        const uint64_t tiles = ((nx + kCarryX - 1) / kCarryX) * ((ny + kCarryY - 1) / kCarryY) * nz;
        h->flow_blocks = static_cast<uint32_t>(std::max<uint64_t>((n_pad + kBlock - 1) / kBlock, tiles));
        h->flow_squares = static_cast<double *>(zeros(h, h->flow_blocks * sizeof(double)));
        h->flow_total = static_cast<double *>(zeros(h, sizeof(double)));
        for (uint64_t i = 0; i < n; ++i) h->air_cells += L.types[i] != 0;
        h->pairs = 0;
    }
    h->layer_state = zeros(h, 3 * L.layer_cells * e);
    if (L.layer_cells > 0) {
        h->layer_face[0] = zeros(h, 3 * n_pad * e);
        h->layer_face[1] = zeros(h, 3 * n_pad * e);
    }
    if (clock.on) check(hipDeviceSynchronize(), "set-up copies");  // so that the copies count
    clock.mark("device copies");
}

template <typename T>
StepArgs<T> step_args(const as_fdtd *h) {
    StepArgs<T> s{};
    s.types = h->types;
    s.scale = h->scale;
    s.tiles = h->tiles;
    s.tile_lanes = h->tile_lanes;
    s.interior_bits = h->interior_bits;
    s.tile_blocks = h->tile_count;
    s.tiles_x = h->tiles_x;
    s.tiles_y = h->tiles_y;
    s.nx = h->nx;
    s.ny = h->ny;
    s.nz = h->nz;
    s.px = h->px;
    s.tile_z = h->tile_z;
    s.sxy = h->sxy;
    s.axial = static_cast<T>(static_cast<float>(h->axial));
    s.diagonal = static_cast<T>(static_cast<float>(h->diagonal));
    s.centre = static_cast<T>(static_cast<float>(h->centre));
    s.total = static_cast<T>(static_cast<float>(h->total));
    s.specials = h->specials;
    // Whole multiples of the dies, so the interior's blocks start on die 0 (see the tiles).
    s.special_blocks = (h->specials + kBlock - 1) / kBlock;
    s.special_blocks = (s.special_blocks + kXcds - 1) / kXcds * kXcds;
    s.sp_cell = h->sp_cell;
    s.sp_links = h->sp_links;
    s.sp_states = h->sp_states;
    s.link_base = h->link_base;
    s.state_base = h->state_base;
    s.link_dirs = h->link_dirs;
    s.link_weight = h->link_weight;
    s.coef_a = h->coef_a;
    s.coef_b = static_cast<const T *>(h->coef_b);
    s.coef_inv = static_cast<const T *>(h->coef_inv);
    s.source = h->source;
    s.state_weight = h->state_weight;
    s.state_branch = h->state_branch;
    s.sp_rec = h->sp_rec;
    s.kind_coef_a = h->kind_coef_a;
    s.kind_coef_b = static_cast<const T *>(h->kind_coef_b);
    s.kind_coef_inv = static_cast<const T *>(h->kind_coef_inv);
    s.kind_source = h->kind_source;
    s.kind_weight = h->kind_weight;
    s.kind_dirs = h->kind_dirs;
    s.kind_state_weight = h->kind_state_weight;
    s.kind_state_branch = h->kind_state_branch;
    s.kind_coef4 = h->kind_coef4;
    s.kind_state4 = h->kind_state4;
    s.kind_links = h->kind_links;
    s.kind_states = h->kind_states;
    s.branch_now = h->branch_now;
    s.branch_before = h->branch_before;
    s.branch_gain = static_cast<const T *>(h->branch_gain);
    s.face_tile = h->face_blocks ? h->face_tile_none : h->face_tile;
    s.face_count = h->face_blocks ? h->face_specials : 0;
    s.face_blocks = (s.face_count + kBlock - 1) / kBlock;
    s.face_blocks = (s.face_blocks + kXcds - 1) / kXcds * kXcds;
    s.face_record = h->face_record;
    s.face_state_kind = h->face_state_kind;
    s.face_coef = h->face_coef;
    s.face_branch = h->face_branch;
    s.face_lookup = h->face_lookup;
    s.layer_cells = h->layer_cells;
    s.layer_blocks = (h->layer_cells + kBlock - 1) / kBlock;
    s.layer_cell = h->layer_cell;
    s.px_magic = h->px_magic;
    s.px_shift = h->px_shift;
    s.ny_magic = h->ny_magic;
    s.ny_shift = h->ny_shift;
    s.layer_state = static_cast<T *>(h->layer_state);
    s.layer_edge = h->layer_edge;
    s.layer_scale = h->layer_scale;
    for (int k = 0; k < 6; ++k) {
        s.keep[k] = h->keep[k];
        s.push[k] = h->push[k];
    }
    s.probes = h->probes;
    s.probe_blocks = (h->probes + kProbesPerBlock - 1) / kProbesPerBlock;
    s.probe_blocks = (s.probe_blocks + kXcds - 1) / kXcds * kXcds;
    s.probe_start = h->probe_start;
    s.probe_cell = h->probe_cell;
    s.probe_weight = h->probe_weight;
    return s;
}

// ---- Moving air: the carrying (fdtd.hpp:1937-2120) ----------------------------------
//
// Before each step both pressures, the current and the previous, are carried along with
// the air by the three-stage strong-stability-preserving Runge-Kutta method, a launch a
// stage (each stage reads only the stage before, so the grid's order does not matter):
//   t1 = q + d(q),  t2 = 3/4 q + 1/4 (t1 + d(t1)),  q = q + 2/3 ((t2 + d(t2)) - q),
// d(q) = -v . grad q over a step at each cell the air moves, cells it does not keeping their
// values to the bit. Plain: per axis, the row's differences as its mask allows (the
// fourth-order central one with two moving cells either side, else half the differences
// across the faces and a twelfth of each loop of three, the cell's own value out and
// silence in through a face to air that does not move), times the cell's shift. Round:
// per axis, the two faces' winds times the differences across them, halved, and the three
// loops; at the border of the moving air, through a face to still air the cell's value goes
// out and silence comes in. The arithmetic, float with no contraction, is the reference's
// in its order.
struct Carry {
    const uint8_t *__restrict__ types;
    const uint8_t *__restrict__ mask;
    const float *__restrict__ fx;
    const float *__restrict__ fy;
    const float *__restrict__ wind[3];
    const float *__restrict__ loop[3];
    float *t1, *t2;
    double *squares;  // a sum per block (the first stage of the current pressure only)
    uint32_t n;
    int px;
    long long sxy;
};

// q's change along a row (stride s) at i over a cell, as air moving by `shift` along it
// carries it; m: which of the row's cells one (bits 0, 1) and two (bits 2, 3) below and
// above move with it.
__device__ __forceinline__ float row_gradient(const float *q, long long i, long long s, unsigned m, float shift) {
    if ((m & 15u) == 15u) return (8.0f * (q[i + s] - q[i - s]) - (q[i + 2 * s] - q[i - 2 * s])) * (1.0f / 12.0f);
    const bool below = m & 1u, above = m & 2u;
    const float c = q[i];
    float faces = 0.0f, loops = 0.0f;
    if (above) {
        faces += 0.5f * (q[i + s] - c);
        if (m & 8u) loops += q[i + s] - q[i + 2 * s];
    } else if (shift < 0.0f) {
        faces -= c;
    }
    if (below) {
        faces += 0.5f * (c - q[i - s]);
        if (m & 4u) loops += q[i - 2 * s] - q[i - s];
    } else if (shift > 0.0f) {
        faces += c;
    }
    if (below && above) loops += q[i + s] - q[i - s];
    return faces + loops * (1.0f / 12.0f);
}

__device__ __forceinline__ float carried_plain(const Carry &c, const float *q, long long i) {
    const float sx = c.fx[i], sy = c.fy[i];
    const unsigned m = c.mask[i];
    return -(sx * row_gradient(q, i, 1, m & 15u, sx) + sy * row_gradient(q, i, c.px, m >> 4, sy));
}

__device__ __forceinline__ float carried_round(const Carry &c, const float *q, long long i) {
    const long long stride[3] = {1, c.px, c.sxy};
    const float v = q[i];
    float sum = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long s = stride[d];
        const float *w = c.wind[d], *l = c.loop[d];
        const float above = q[i + s], below = q[i - s];
        sum += 0.5f * (w[i] * (above - v) + w[i - s] * (v - below)) + l[i] * (above - q[i + 2 * s]) +
               l[i - s] * (above - below) + l[i - 2 * s] * (q[i - 2 * s] - below);
    }
    return -sum;
}

__device__ __forceinline__ float carried_border(const Carry &c, const float *q, long long i) {
    const long long stride[3] = {1, c.px, c.sxy};
    const float v = q[i];
    float sum = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long long s = stride[d];
        const float *w = c.wind[d], *l = c.loop[d];
        const float above = q[i + s], below = q[i - s];
        if (c.mask[i + s]) {
            sum += 0.5f * w[i] * (above - v);
        } else if (w[i] < 0.0f) {
            sum -= w[i] * v;
        }
        if (c.mask[i - s]) {
            sum += 0.5f * w[i - s] * (v - below);
        } else if (w[i - s] > 0.0f) {
            sum += w[i - s] * v;
        }
        sum += l[i] * (above - q[i + 2 * s]) + l[i - s] * (above - below) + l[i - 2 * s] * (q[i - 2 * s] - below);
    }
    return -sum;
}

// Stage STAGE (0, 1, 2) of carrying q: each cell off the solid from the stage before. With
// SQUARES (the first stage of the current pressure), q's squares over the air as it comes
// (fdtd.hpp:2069-2072) are summed over the block in double, in a fixed order, and added to
// the block's own slot: the sums are the same run after run.
template <bool ROUND, int STAGE, bool SQUARES>
__global__ __launch_bounds__(kBlock) void carry_kernel(const Carry c, float *q) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool air = i < c.n && c.types[i] != 0;
    if constexpr (SQUARES) {
        __shared__ double part[kBlock];
        const double v = air ? static_cast<double>(q[i]) : 0.0;
        part[threadIdx.x] = v * v;
        __syncthreads();
#pragma unroll
        for (unsigned w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) c.squares[blockIdx.x] += part[0];
    }
    if (!air) return;
    const bool still = ROUND ? c.mask[i] == 0 : (c.fx[i] == 0.0f && c.fy[i] == 0.0f);
    if (STAGE == 2 && still) return;
    const float *from = STAGE == 0 ? q : (STAGE == 1 ? c.t1 : c.t2);
    float d = 0.0f;
    if (!still) {
        if (ROUND) {
            d = c.mask[i] == 2 ? carried_border(c, from, i) : carried_round(c, from, i);
        } else {
            d = carried_plain(c, from, i);
        }
    }
    if (STAGE == 0) {
        c.t1[i] = still ? q[i] : q[i] + d;
    } else if (STAGE == 1) {
        c.t2[i] = still ? q[i] : 0.75f * q[i] + 0.25f * (c.t1[i] + d);
    } else {
        q[i] = q[i] + (2.0f / 3.0f) * ((c.t2[i] + d) - q[i]);
    }
}

Carry carry_args(const as_fdtd *h) {
    Carry c{};
    c.types = h->types;
    c.mask = h->flow_mask;
    c.fx = h->flow_x;
    c.fy = h->flow_y;
    for (int d = 0; d < 3; ++d) {
        c.wind[d] = h->flow_wind[d];
        c.loop[d] = h->flow_loop[d];
    }
    c.t1 = h->flow_t1;
    c.t2 = h->flow_t2;
    c.squares = h->flow_squares;
    c.n = static_cast<uint32_t>(h->n_pad);
    c.px = h->px;
    c.sxy = h->sxy;
    return c;
}

// The three stages of carrying q, on `stream`; `squares`: q's squares summed on the way.
template <bool ROUND>
void carry_stages(const Carry &c, float *q, bool squares, hipStream_t stream) {
    const dim3 grid((c.n + kBlock - 1) / kBlock), block(kBlock);
    if (squares) {
        hipLaunchKernelGGL((carry_kernel<ROUND, 0, true>), grid, block, 0, stream, c, q);
    } else {
        hipLaunchKernelGGL((carry_kernel<ROUND, 0, false>), grid, block, 0, stream, c, q);
    }
    hipLaunchKernelGGL((carry_kernel<ROUND, 1, false>), grid, block, 0, stream, c, q);
    hipLaunchKernelGGL((carry_kernel<ROUND, 2, false>), grid, block, 0, stream, c, q);
}

void carry(const as_fdtd *h, float *q, bool squares, hipStream_t stream) {
    const Carry c = carry_args(h);
    if (h->flow_round) {
        carry_stages<true>(c, q, squares, stream);
    } else {
        carry_stages<false>(c, q, squares, stream);
    }
}

// Plain rows, both pressures a step in one launch. The air moves a cell along its row in x and
// in y only, so each plane is carried on its own. A block takes a tile of kCarryX x kCarryY
// cells of one plane and its frame: the tile and kCarryHalo cells round it, two for each
// stage's reach. The frame's pressures, shifts, masks and kinds go to LDS; the first stage is
// taken over the frame less two cells a side, the second less four, the third over the tile,
// each from the stage before in LDS. A halo cell is carried by every block whose frame holds
// it, by the same operations, so each value equals the one-stage-a-launch carrying's.
struct PlainCarry {
    const uint8_t *__restrict__ types;
    const uint8_t *__restrict__ mask;
    const float *__restrict__ fx;
    const float *__restrict__ fy;
    float *a, *b;     // p(k - 1) and p(k), carried in place
    double *squares;  // a sum a block: b's squares over the tile's air as they come
    uint32_t nx, ny, tiles_x, tiles_y;
    int px;
    long long sxy;
};

__global__ __launch_bounds__(kBlock) void plain_carry_kernel(const PlainCarry c) {
    __shared__ float q[kFrame], t1[kFrame], t2[kFrame], sx[kFrame], sy[kFrame];
    __shared__ uint8_t m[kFrame], kind[kFrame];  // kind: 0 solid or off the grid, 1 still, 2 moved
    __shared__ double part[kBlock];
    const uint32_t per_plane = c.tiles_x * c.tiles_y;
    const uint32_t z = blockIdx.x / per_plane, r0 = blockIdx.x % per_plane;
    const int x0 = static_cast<int>(r0 % c.tiles_x) * kCarryX - kCarryHalo;
    const int y0 = static_cast<int>(r0 / c.tiles_x) * kCarryY - kCarryHalo;
    const long long base = static_cast<long long>(z) * c.sxy;
    // The frame's cell f, at (x0 + f % kFrameX, y0 + f / kFrameX): its index, -1 off the grid.
    const auto cell = [&](int f) -> long long {
        const int x = x0 + f % kFrameX, y = y0 + f / kFrameX;
        if (x < 0 || y < 0 || x >= static_cast<int>(c.nx) || y >= static_cast<int>(c.ny)) return -1;
        return base + static_cast<long long>(y) * c.px + x;
    };
    for (int f = threadIdx.x; f < kFrame; f += kBlock) {
        const long long i = cell(f);
        uint8_t k = 0, bits = 0;
        float vx = 0.0f, vy = 0.0f;
        if (i >= 0 && c.types[i] != 0) {
            vx = c.fx[i];
            vy = c.fy[i];
            bits = c.mask[i];
            k = vx == 0.0f && vy == 0.0f ? 1 : 2;
        }
        kind[f] = k;
        m[f] = bits;
        sx[f] = vx;
        sy[f] = vy;
    }
    // -v . grad p at frame cell f, over a step (carried_plain's operations).
    const auto carried = [&](const float *p, int f) {
        const unsigned bits = m[f];
        return -(sx[f] * row_gradient(p, f, 1, bits & 15u, sx[f]) + sy[f] * row_gradient(p, f, kFrameX, bits >> 4, sy[f]));
    };
    for (int field = 0; field < 2; ++field) {
        float *p = field == 0 ? c.b : c.a;
        __syncthreads();  // the frame's kinds are in; the field before has left q, t1, t2
        for (int f = threadIdx.x; f < kFrame; f += kBlock) {
            const long long i = cell(f);
            q[f] = i >= 0 ? p[i] : 0.0f;
        }
        __syncthreads();
        if (field == 0) {
            double sum = 0.0;
            for (int r = threadIdx.x; r < kCarryX * kCarryY; r += kBlock) {
                const int f = (kCarryHalo + r / kCarryX) * kFrameX + kCarryHalo + r % kCarryX;
                if (kind[f] != 0) sum += static_cast<double>(q[f]) * static_cast<double>(q[f]);
            }
            part[threadIdx.x] = sum;
            __syncthreads();
#pragma unroll
            for (unsigned w = kBlock / 2; w > 0; w >>= 1) {
                if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) c.squares[blockIdx.x] += part[0];
        }
        constexpr int w1 = kFrameX - 4, h1 = kFrameY - 4;
        for (int r = threadIdx.x; r < w1 * h1; r += kBlock) {
            const int f = (2 + r / w1) * kFrameX + 2 + r % w1;
            if (kind[f] == 0) continue;
            t1[f] = kind[f] == 1 ? q[f] : q[f] + carried(q, f);
        }
        __syncthreads();
        constexpr int w2 = kFrameX - 8, h2 = kFrameY - 8;
        for (int r = threadIdx.x; r < w2 * h2; r += kBlock) {
            const int f = (4 + r / w2) * kFrameX + 4 + r % w2;
            if (kind[f] == 0) continue;
            t2[f] = kind[f] == 1 ? q[f] : 0.75f * q[f] + 0.25f * (t1[f] + carried(t1, f));
        }
        __syncthreads();
        for (int r = threadIdx.x; r < kCarryX * kCarryY; r += kBlock) {
            const int f = (kCarryHalo + r / kCarryX) * kFrameX + kCarryHalo + r % kCarryX;
            if (kind[f] != 2) continue;
            p[cell(f)] = q[f] + (2.0f / 3.0f) * ((t2[f] + carried(t2, f)) - q[f]);
        }
    }
}

// Both pressures carried a step: plain rows by tiles of planes in one launch (unless the
// launch shape is split), else three launches of a stage each per pressure; b's squares
// summed on the way.
void carry_both(const as_fdtd *h, float *a, float *b, hipStream_t stream) {
    if (h->flow_round || h->split) {
        carry(h, b, true, stream);
        carry(h, a, false, stream);
        return;
    }
    PlainCarry c{};
    c.types = h->types;
    c.mask = h->flow_mask;
    c.fx = h->flow_x;
    c.fy = h->flow_y;
    c.a = a;
    c.b = b;
    c.squares = h->flow_squares;
    c.nx = h->nx;
    c.ny = h->ny;
    c.tiles_x = (h->nx + kCarryX - 1) / kCarryX;
    c.tiles_y = (h->ny + kCarryY - 1) / kCarryY;
    c.px = h->px;
    c.sxy = h->sxy;
    hipLaunchKernelGGL(plain_carry_kernel, dim3(c.tiles_x * c.tiles_y * h->nz), dim3(kBlock), 0, stream, c);
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double *p, uint32_t count) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// The blocks' sums into one, in a fixed order: each lane its strided share in turn, then
// the lanes pairwise. One block.
__global__ __launch_bounds__(kBlock) void total_kernel(const double *__restrict__ parts, uint32_t count,
                                                       double *__restrict__ total) {
    __shared__ double lane[kBlock];
    double sum = 0.0;
    for (uint32_t k = threadIdx.x; k < count; k += kBlock) sum += parts[k];
    lane[threadIdx.x] = sum;
    __syncthreads();
#pragma unroll
    for (unsigned w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) lane[threadIdx.x] += lane[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = lane[0];
}

// A run of `steps` starts: in moving air its squares start from zero (by a kernel, so that a
// caller may capture it).
void begin_run(as_fdtd *h, size_t steps) {
    if (!h->flowing) return;
    hipLaunchKernelGGL(fill_kernel, dim3((h->flow_blocks + kBlock - 1) / kBlock), dim3(kBlock), 0, h->stream,
                       h->flow_squares, h->flow_blocks);
    h->run_steps = steps;
}

// A step as one launch of `blocks` blocks (the roles in the order step_kernel reads them).
template <typename T>
void launch_fused(const as_fdtd *h, const StepArgs<T> &s, uint32_t blocks, hipStream_t stream) {
    const dim3 grid(blocks), block(kBlock);
    if (s.sp_rec != nullptr) {
        if (h->wide && h->scaled) {
            hipLaunchKernelGGL((step_kernel<T, true, true, true>), grid, block, 0, stream, s);
        } else if (h->wide) {
            hipLaunchKernelGGL((step_kernel<T, true, false, true>), grid, block, 0, stream, s);
        } else if (h->scaled) {
            hipLaunchKernelGGL((step_kernel<T, false, true, true>), grid, block, 0, stream, s);
        } else {
            hipLaunchKernelGGL((step_kernel<T, false, false, true>), grid, block, 0, stream, s);
        }
    } else if (h->wide && h->scaled) {
        hipLaunchKernelGGL((step_kernel<T, true, true, false>), grid, block, 0, stream, s);
    } else if (h->wide) {
        hipLaunchKernelGGL((step_kernel<T, true, false, false>), grid, block, 0, stream, s);
    } else if (h->scaled) {
        hipLaunchKernelGGL((step_kernel<T, false, true, false>), grid, block, 0, stream, s);
    } else {
        hipLaunchKernelGGL((step_kernel<T, false, false, false>), grid, block, 0, stream, s);
    }
}

// A step as a launch per kind of cell (the split launch shape); the probes last, where asked.
template <typename T>
void launch_split(const as_fdtd *h, const StepArgs<T> &s, hipStream_t stream, bool probes) {
    if (s.special_blocks && s.sp_rec != nullptr)
        hipLaunchKernelGGL((special_kernel<T, true>), dim3(s.special_blocks), dim3(kBlock), 0, stream, s);
    else if (s.special_blocks)
        hipLaunchKernelGGL((special_kernel<T, false>), dim3(s.special_blocks), dim3(kBlock), 0, stream, s);
    if (s.face_blocks) hipLaunchKernelGGL((face_kernel<T>), dim3(s.face_blocks), dim3(kBlock), 0, stream, s);
    if (s.tile_blocks) {
        const dim3 grid(s.tile_blocks), block(kBlock);
        if (h->wide && h->scaled) {
            hipLaunchKernelGGL((interior_kernel<T, true, true>), grid, block, 0, stream, s);
        } else if (h->wide) {
            hipLaunchKernelGGL((interior_kernel<T, true, false>), grid, block, 0, stream, s);
        } else if (h->scaled) {
            hipLaunchKernelGGL((interior_kernel<T, false, true>), grid, block, 0, stream, s);
        } else {
            hipLaunchKernelGGL((interior_kernel<T, false, false>), grid, block, 0, stream, s);
        }
    }
    if (s.layer_blocks) hipLaunchKernelGGL((layer_kernel<T>), dim3(s.layer_blocks), dim3(kBlock), 0, stream, s);
    const uint32_t probe_blocks = (h->probes + kProbesPerBlock - 1) / kProbesPerBlock;
    if (probes && probe_blocks) hipLaunchKernelGGL((probe_kernel<T>), dim3(probe_blocks), dim3(kBlock), 0, stream, s);
}

// One step from the fields in slots prev (p(k - 1), overwritten by p(k + 1)) and prev ^ 1
// (p(k)): the reference's swaps (fdtd.hpp:883-885, 900-903) are the slots' exchange.
template <typename T>
void launch_one(as_fdtd *h, StepArgs<T> &s, uint32_t blocks, int prev, unsigned long long k, hipStream_t stream) {
    s.a = static_cast<const T *>(h->field[prev]);
    s.o = static_cast<T *>(h->field[prev]);  // in place
    s.b = static_cast<const T *>(h->field[prev ^ 1]);
    s.wa = static_cast<const T *>(h->velocity[prev]);
    s.wo = static_cast<T *>(h->velocity[prev]);
    s.wb = static_cast<const T *>(h->velocity[prev ^ 1]);
    s.layer_face_in = static_cast<const T *>(h->layer_face[prev & 1]);
    s.layer_face_out = static_cast<T *>(h->layer_face[(prev & 1) ^ 1]);
    s.step = k;
    if constexpr (std::is_same<T, float>::value) {
        if (h->flowing) {
            // The probes read p(k) as it stands, then both pressures are carried with the
            // air (fdtd.hpp:838-846), then the step, without its probe blocks.
            const uint32_t probe_blocks = (h->probes + kProbesPerBlock - 1) / kProbesPerBlock;
            if (probe_blocks) hipLaunchKernelGGL((probe_kernel<T>), dim3(probe_blocks), dim3(kBlock), 0, stream, s);
            carry_both(h, static_cast<float *>(h->field[prev]), static_cast<float *>(h->field[prev ^ 1]), stream);
            StepArgs<T> t = s;
            const uint32_t rest = blocks - t.probe_blocks;
            t.probe_blocks = 0;
            t.probes = 0;
            if (h->split) {
                launch_split<T>(h, t, stream, false);
                return;
            }
            if (rest > 0) launch_fused<T>(h, t, rest, stream);
            return;
        }
    }
    if (h->split) {
        launch_split<T>(h, s, stream, true);
        return;
    }
    launch_fused<T>(h, s, blocks, stream);
}

template <typename T>
uint32_t block_count(const as_fdtd *h, const StepArgs<T> &s) {
    (void)h;
    return s.probe_blocks + s.special_blocks + s.face_blocks + s.tile_blocks + s.layer_blocks + s.cell_blocks;
}

// The boxes' arguments (pair_kernel), their fields set per launch.
BoxArgs box_args(const as_fdtd *h) {
    BoxArgs x{};
    x.bits = h->box_bits;
    x.boxes = h->boxes;
    x.box_blocks = h->box_blocks;
    x.boxes_x = h->boxes_x;
    x.groups_y = h->groups_y;
    x.ny = h->ny;
    x.nz = h->nz;
    x.px = h->px;
    x.row_bytes = static_cast<uint32_t>(h->px) * 4u;
    x.plane_bytes = static_cast<uint32_t>(h->sxy) * 4u;
    x.field_bytes = static_cast<uint32_t>(h->n_pad * 4u);
    x.bits_bytes = static_cast<uint32_t>(static_cast<uint64_t>(h->nz) * h->groups_y * h->px * 4u);
    x.centre = static_cast<float>(h->centre);
    x.axial = static_cast<float>(h->axial);
    return x;
}

// Steps k0 to k1 - 1 from slot `prev`: two steps by two launches where the handle has them
// (see pair_kernel), the rest one step a launch. Returns the slot of p(k1 - 1).
template <typename T>
int enqueue_steps(as_fdtd *h, StepArgs<T> s, uint32_t blocks, int prev, size_t k0, size_t k1, hipStream_t stream) {
    size_t k = k0;
    if constexpr (std::is_same<T, float>::value) {
        if (h->pairs && k + 2 <= k1) {
            BoxArgs x = box_args(h);
            // Both launches step the face specials by their own blocks; the second steps the
            // listed interior cells where one step a launch has its tiles.
            StepArgs<float> first = s, second = s;
            for (StepArgs<float> *t : {&first, &second}) {
                t->tile_blocks = 0;
                t->face_tile = h->face_tile_none;
                t->face_count = h->face_specials;
                t->face_blocks = (h->face_specials + kBlock - 1) / kBlock;
                t->face_blocks = (t->face_blocks + kXcds - 1) / kXcds * kXcds;
            }
            second.cells = h->cells;
            second.cell_count = h->cell_count;
            second.cell_blocks = (h->cell_count + kBlock - 1) / kBlock;
            if ((x.box_blocks > 0 && (x.boxes == nullptr || x.bits == nullptr)) ||
                (second.cell_blocks > 0 && second.cells == nullptr) ||
                (first.face_blocks > 0 && first.face_record == nullptr)) {
                throw Failure{AS_EINVAL, "internal: a two-step launch lacks its arrays"};
            }
            const uint32_t first_blocks =
                first.probe_blocks + first.special_blocks + first.face_blocks + x.box_blocks + first.layer_blocks;
            const uint32_t second_blocks = block_count(h, second);
            for (; k + 2 <= k1; k += 2) {
                // p(k + 1) into slot next, p(k + 2) into next ^ 1. A pair keeps prev & 1, the
                // slot of the face velocities, which it writes and then overwrites.
                const int next = (prev < 2 ? 2 : 0) | (prev & 1);
                float *f[4] = {static_cast<float *>(h->field[prev]), static_cast<float *>(h->field[prev ^ 1]),
                               static_cast<float *>(h->field[next]), static_cast<float *>(h->field[next ^ 1])};
                float *v[4] = {static_cast<float *>(h->velocity[prev]), static_cast<float *>(h->velocity[prev ^ 1]),
                               static_cast<float *>(h->velocity[next]), static_cast<float *>(h->velocity[next ^ 1])};
                first.a = f[0];
                first.b = f[1];
                first.o = f[2];
                first.wa = v[0];
                first.wb = v[1];
                first.wo = v[2];
                first.layer_face_in = static_cast<const float *>(h->layer_face[prev & 1]);
                first.layer_face_out = static_cast<float *>(h->layer_face[(prev & 1) ^ 1]);
                first.step = k;
                x.prev = f[0];
                x.cur = f[1];
                x.out1 = f[2];
                x.out2 = f[3];
                if (first_blocks) {
                    if (h->pair_z == 16) {
                        hipLaunchKernelGGL((pair_kernel<16>), dim3(first_blocks), dim3(kBlock), 0, stream, first, x);
                    } else {
                        hipLaunchKernelGGL((pair_kernel<8>), dim3(first_blocks), dim3(kBlock), 0, stream, first, x);
                    }
                }
                second.a = f[1];
                second.b = f[2];
                second.o = f[3];
                second.wa = v[1];
                second.wb = v[2];
                second.wo = v[3];
                second.layer_face_in = static_cast<const float *>(h->layer_face[(prev & 1) ^ 1]);
                second.layer_face_out = static_cast<float *>(h->layer_face[prev & 1]);
                second.step = k + 1;
                if (second_blocks && second.sp_rec != nullptr) {
                    hipLaunchKernelGGL((step_kernel<float, false, false, true>), dim3(second_blocks), dim3(kBlock), 0,
                                       stream, second);
                } else if (second_blocks) {
                    hipLaunchKernelGGL((step_kernel<float, false, false, false>), dim3(second_blocks), dim3(kBlock), 0,
                                       stream, second);
                }
                prev = next;
            }
        }
    }
    for (; k < k1; ++k) {
        launch_one<T>(h, s, blocks, prev, k, stream);
        prev ^= 1;
    }
    return prev;
}

// Graphs of graph_steps steps, one for each slot the run may start from (built at need),
// on staging buffers for the drive and the probes' output: a run replays them, copying
// its drive in and the probes out between replays, so the host launches one graph where
// it would launch graph_steps kernels.
template <typename T>
void ensure_graph(as_fdtd *h, int prev) {
    if (h->graph_exec[prev] != nullptr) return;
    const uint32_t G = h->graph_steps;
    if (h->sig_stage == nullptr) {
        check(hipMalloc(reinterpret_cast<void **>(&h->sig_stage), G * sizeof(float)), "hipMalloc");
        check(hipMemset(h->sig_stage, 0, G * sizeof(float)), "hipMemset");
        const size_t out_len = static_cast<size_t>(std::max<uint32_t>(h->probes, 1)) * G;
        check(hipMalloc(reinterpret_cast<void **>(&h->out_stage), out_len * sizeof(double)), "hipMalloc");
    }
    StepArgs<T> s = step_args<T>(h);
    s.signal = h->sig_stage;
    s.out = h->out_stage;
    s.out_stride = G;
    const uint32_t blocks = block_count(h, s);
    capture_graph(h->stream, &h->graph_exec[prev],
                  [&] { h->graph_end[prev] = enqueue_steps<T>(h, s, blocks, prev, 0, G, h->stream); });
}

template <typename T>
void launch_steps(as_fdtd *h, const float *d_signal, size_t steps, double *d_out, size_t out_stride) {
    StepArgs<T> s = step_args<T>(h);
    const uint32_t blocks = block_count(h, s);
    if (blocks == 0) return;
    size_t k = 0;
    int prev = h->prev;
    const uint32_t G = h->graph_steps;
    if (G > 0 && steps >= G) {
        for (; k + G <= steps; k += G) {
            ensure_graph<T>(h, prev);
            check(hipMemcpyAsync(h->sig_stage, d_signal + k, G * sizeof(float), hipMemcpyDeviceToDevice, h->stream),
                  "hipMemcpyAsync");
            check(hipGraphLaunch(h->graph_exec[prev], h->stream), "hipGraphLaunch");
            prev = h->graph_end[prev];
            if (h->probes > 0) {
                check(hipMemcpy2DAsync(d_out + k, out_stride * sizeof(double), h->out_stage, G * sizeof(double),
                                       G * sizeof(double), h->probes, hipMemcpyDeviceToDevice, h->stream),
                      "hipMemcpy2DAsync");
            }
        }
    }
    s.signal = d_signal;
    s.out = d_out;
    s.out_stride = out_stride;
    prev = enqueue_steps<T>(h, s, blocks, prev, k, steps, h->stream);
    check(hipGetLastError(), "step launch");
    h->prev = prev;
    h->done += steps;
}

int fail(as_fdtd *h, const Failure &f) {
    if (h) h->error = f.what;
    return f.code;
}

}  // namespace

// The no-allocation run (as_fdtd_run_static) stages each piece of a run in this array: the
// probes' output (probes x piece doubles) from its start, the drive (piece floats) from the
// next 256-byte boundary after it. Each device holds its own copy, at its own address, for
// the life of the process. In host code `scratch` names a host-side stand-in, not device
// memory: its device address comes from hipGetSymbolAddress (scratch_address).
constexpr size_t kScratchBytes = size_t{1} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

struct ScratchAddress {
    std::once_flag once;
    void *ptr = nullptr;
    hipError_t error = hipSuccess;
};
ScratchAddress g_scratch[kMaxDevices];

// The device address of `scratch` on `device`, looked up once per device with that device
// current, and the caller's current device restored.
void *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) refuse("the handle's device lies beyond the library's table");
    ScratchAddress &slot = g_scratch[device];
    std::call_once(slot.once, [&] {
        int current = -1;
        slot.error = hipGetDevice(&current);
        if (slot.error != hipSuccess) return;
        if (current != device) {
            slot.error = hipSetDevice(device);
            if (slot.error != hipSuccess) return;
        }
        slot.error = hipGetSymbolAddress(&slot.ptr, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
    });
    check(slot.error, "hipGetSymbolAddress(scratch)");
    if (slot.ptr == nullptr) throw Failure{AS_EDEVICE, "hipGetSymbolAddress(scratch): no address"};
    return slot.ptr;
}

// `steps` steps launched one a step from the host (no graphs: capturing and instantiating
// one allocates), on device pointers.
template <typename T>
void launch_direct(as_fdtd *h, const float *d_signal, size_t steps, double *d_out, size_t out_stride) {
    StepArgs<T> s = step_args<T>(h);
    const uint32_t blocks = block_count(h, s);
    if (blocks == 0) return;
    s.signal = d_signal;
    s.out = d_out;
    s.out_stride = out_stride;
    h->prev = enqueue_steps<T>(h, s, blocks, h->prev, 0, steps, h->stream);
    check(hipGetLastError(), "step launch");
    h->done += steps;
}

size_t round_up_256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" {

as_fdtd *as_fdtd_create(const as_layout *layout, int precision, const as_tuning *tuning) {
    g_create_error.clear();
    if (layout == nullptr) {
        g_create_error = "no layout";
        return nullptr;
    }
    as_fdtd *h = new (std::nothrow) as_fdtd();
    if (h == nullptr) {
        g_create_error = "out of host memory";
        return nullptr;
    }
    try {
        build(h, *layout, precision, tuning);
        return h;
    } catch (const Failure &f) {
        g_create_error = f.what;
    } catch (const std::exception &e) {
        g_create_error = e.what();
    }
    as_fdtd_destroy(h);
    return nullptr;
}

const char *as_fdtd_create_error(void) { return g_create_error.c_str(); }

int as_fdtd_run_device(as_fdtd *h, const float *d_signal, size_t steps, double *d_out) {
    if (h == nullptr) return AS_EINVAL;
    try {
        if (steps == 0) return AS_OK;
        begin_run(h, steps);
        if (h->precision == 64) {
            launch_steps<double>(h, d_signal, steps, d_out, steps);
        } else {
            launch_steps<float>(h, d_signal, steps, d_out, steps);
        }
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_run(as_fdtd *h, const float *signal, size_t steps, double *out) {
    if (h == nullptr) return AS_EINVAL;
    try {
        if (steps == 0) return AS_OK;
        if (signal == nullptr || (out == nullptr && h->probes > 0)) refuse("run needs a signal and room for the probes");
        // Pieces of at most this many steps, so the device's output stays small.
        const size_t piece = 1u << 16;
        if (h->signal_cap < std::min(steps, piece)) {
            if (h->signal_buf) (void)hipFree(h->signal_buf);
            h->signal_buf = nullptr;
            h->signal_cap = std::min(steps, piece);
            check(hipMalloc(reinterpret_cast<void **>(&h->signal_buf), h->signal_cap * sizeof(float)), "hipMalloc");
        }
        const size_t out_need = std::max<size_t>(1, h->probes) * std::min(steps, piece);
        if (h->out_cap < out_need) {
            if (h->out_buf) (void)hipFree(h->out_buf);
            h->out_buf = nullptr;
            h->out_cap = out_need;
            check(hipMalloc(reinterpret_cast<void **>(&h->out_buf), h->out_cap * sizeof(double)), "hipMalloc");
        }
        std::vector<double> chunk;
        begin_run(h, steps);
        for (size_t done = 0; done < steps; done += piece) {
            const size_t m = std::min(piece, steps - done);
            check(hipMemcpyAsync(h->signal_buf, signal + done, m * sizeof(float), hipMemcpyHostToDevice, h->stream),
                  "hipMemcpyAsync");
            if (h->precision == 64) {
                launch_steps<double>(h, h->signal_buf, m, h->out_buf, m);
            } else {
                launch_steps<float>(h, h->signal_buf, m, h->out_buf, m);
            }
            chunk.resize(static_cast<size_t>(h->probes) * m);
            if (!chunk.empty()) {
                check(hipMemcpyAsync(chunk.data(), h->out_buf, chunk.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     h->stream),
                      "hipMemcpyAsync");
            }
            check(hipStreamSynchronize(h->stream), "step");
            for (uint32_t p = 0; p < h->probes; ++p) {
                std::memcpy(out + static_cast<size_t>(p) * steps + done, chunk.data() + static_cast<size_t>(p) * m,
                            m * sizeof(double));
            }
        }
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_run_static(as_fdtd *h, const float *signal, size_t steps, double *out) {
    if (h == nullptr) return AS_EINVAL;
    try {
        if (steps == 0) return AS_OK;
        if (signal == nullptr || (out == nullptr && h->probes > 0)) refuse("run needs a signal and room for the probes");
        const size_t probe_bytes = sizeof(double) * h->probes;
        size_t piece = kScratchBytes / (probe_bytes + sizeof(float));
        while (piece > 0 && round_up_256(piece * probe_bytes) + piece * sizeof(float) > kScratchBytes) --piece;
        if (piece == 0) {
            refuse("as_fdtd_run_static stages a step's probes in " + std::to_string(kScratchBytes) +
                   " bytes: at most " + std::to_string((kScratchBytes - 256) / sizeof(double)) + " probes, not " +
                   std::to_string(h->probes));
        }
        unsigned char *base = static_cast<unsigned char *>(scratch_address(h->device));
        double *d_out = reinterpret_cast<double *>(base);
        float *d_signal = reinterpret_cast<float *>(base + round_up_256(piece * probe_bytes));
        begin_run(h, steps);
        for (size_t done = 0; done < steps; done += piece) {
            const size_t m = std::min(piece, steps - done);
            check(hipMemcpyAsync(d_signal, signal + done, m * sizeof(float), hipMemcpyHostToDevice, h->stream),
                  "hipMemcpyAsync");
            if (h->precision == 64) {
                launch_direct<double>(h, d_signal, m, d_out, m);
            } else {
                launch_direct<float>(h, d_signal, m, d_out, m);
            }
            if (h->probes > 0) {
                check(hipMemcpy2DAsync(out + done, steps * sizeof(double), d_out, m * sizeof(double),
                                       m * sizeof(double), h->probes, hipMemcpyDeviceToHost, h->stream),
                      "hipMemcpy2DAsync");
            }
            // The next piece reuses the staging: this one's copies end first.
            check(hipStreamSynchronize(h->stream), "step");
        }
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_mean_square(as_fdtd *h, double *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    try {
        *out = 0.0;
        if (!h->flowing || h->run_steps == 0 || h->air_cells == 0) return AS_OK;
        hipLaunchKernelGGL(total_kernel, dim3(1), dim3(kBlock), 0, h->stream, h->flow_squares, h->flow_blocks,
                           h->flow_total);
        double total = 0.0;
        check(hipMemcpyAsync(&total, h->flow_total, sizeof(double), hipMemcpyDeviceToHost, h->stream),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(h->stream), "mean square");
        *out = total / (static_cast<double>(h->run_steps) * static_cast<double>(h->air_cells));
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_reset(as_fdtd *h) {
    if (h == nullptr) return AS_EINVAL;
    try {
        const size_t e = h->precision == 64 ? sizeof(double) : sizeof(float);
        for (int k = 0; k < 4; ++k) {
            if (h->field[k] == nullptr) continue;
            check(hipMemsetAsync(h->field[k], 0, h->n_pad * e, h->stream), "hipMemsetAsync");
            check(hipMemsetAsync(h->velocity[k], 0, std::max<size_t>(h->state_slots * e, 256), h->stream),
                  "hipMemsetAsync");
        }
        check(hipMemsetAsync(h->layer_state, 0, std::max<size_t>(3ull * h->layer_cells * e, 256), h->stream),
              "hipMemsetAsync");
        for (void *f : h->layer_face) {
            if (f != nullptr) check(hipMemsetAsync(f, 0, 3 * h->n_pad * e, h->stream), "hipMemsetAsync");
        }
        for (float *t : {h->flow_t1, h->flow_t2}) {
            if (t != nullptr) check(hipMemsetAsync(t, 0, h->n_pad * sizeof(float), h->stream), "hipMemsetAsync");
        }
        check(hipStreamSynchronize(h->stream), "reset");
        h->done = 0;
        h->prev = 0;
        h->run_steps = 0;
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_field(as_fdtd *h, double *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    try {
        const void *b = h->field[h->prev ^ 1];
        const size_t e = h->precision == 64 ? sizeof(double) : sizeof(float);
        std::vector<unsigned char> host(h->n_pad * e);
        check(hipMemcpyAsync(host.data(), b, host.size(), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(h->stream), "field");
        const size_t nx = static_cast<size_t>(h->nx), ny = static_cast<size_t>(h->ny), nz = static_cast<size_t>(h->nz);
        const size_t px = static_cast<size_t>(h->px);
        for (size_t z = 0; z < nz; ++z)
            for (size_t y = 0; y < ny; ++y)
                for (size_t x = 0; x < nx; ++x) {
                    const size_t p = x + px * (y + ny * z);
                    const size_t i = x + nx * (y + ny * z);
                    if (h->precision == 64) {
                        out[i] = reinterpret_cast<const double *>(host.data())[p];
                    } else {
                        out[i] = static_cast<double>(reinterpret_cast<const float *>(host.data())[p]);
                    }
                }
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

}  // extern "C"

namespace {

void free_record(as_fdtd *h) {
    if (h->rec_ckpt) (void)hipFree(h->rec_ckpt);
    if (h->rec_signal) (void)hipFree(h->rec_signal);
    if (h->rec_hist) (void)hipFree(h->rec_hist);
    h->rec_ckpt = nullptr;
    h->rec_signal = nullptr;
    h->rec_hist = nullptr;
    h->rec_steps = h->rec_every = 0;
}

// The record's graphs and their staging (see as_fdtd::rec_graph): graphs first, then the
// buffers their nodes point at.
void free_record_graphs(as_fdtd *h) {
    for (auto &g : h->rec_graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    for (void *p : {static_cast<void *>(h->rec_sig_stage), static_cast<void *>(h->rec_out_stage), h->rec_hist_stage})
        if (p) (void)hipFree(p);
    h->rec_sig_stage = nullptr;
    h->rec_out_stage = nullptr;
    h->rec_hist_stage = nullptr;
    h->rec_graph_every = 0;
}

// The most device memory a record keeps its history in by default (keep_history 0): at
// most this, and at most a quarter of what is free.
constexpr size_t kHistoryBudget = size_t{16} << 30;

// Steps 0 .. count - 1 of a recorded segment, from slot `prev`, one launch a step, s's
// drive, probes' output and history (row k: step k) as given; returns the slot after.
template <typename T>
int record_steps(as_fdtd *h, StepArgs<T> s, uint32_t blocks, int prev, size_t count) {
    for (size_t k = 0; k < count; ++k) {
        launch_one<T>(h, s, blocks, prev, k, h->stream);
        prev ^= 1;
    }
    return prev;
}

void check_adjoint(const as_fdtd *h) {
    if (h->face_specials > 0) refuse("the adjoint needs a handle made with no_faces set");
    if (h->flowing) refuse("the adjoint steps still air only: moving air has no transposed carrying here");
    if (h->layer_cells > 0 && (h->px > (1 << 30) || h->ny > (1 << 30)))
        refuse("the adjoint's matched layer needs a grid under 2^30 cells along each axis");
}

// The forward half of the gradient: from rest, `steps` steps, the state saved before each
// segment of `every` steps, the probes into out (host, probes x steps) if given.
template <typename T>
void record(as_fdtd *h, const float *signal, size_t steps, size_t every, double *out) {
    check_adjoint(h);
    if (every == 0) refuse("checkpoints every 0 steps");
    free_record(h);
    const size_t e = sizeof(T), P = h->probes;
    const size_t field = h->n_pad * e, states = h->state_slots * e;
    const size_t segments = (steps + every - 1) / every;
    const size_t per = 2 * field + 2 * states;
    check(hipMalloc(reinterpret_cast<void **>(&h->rec_ckpt), std::max<size_t>(segments * per, 256)), "hipMalloc");
    check(hipMalloc(reinterpret_cast<void **>(&h->rec_signal), steps * sizeof(float)), "hipMalloc");
    double *d_out = nullptr;
    check(hipMalloc(reinterpret_cast<void **>(&d_out), std::max<size_t>(P, 1) * steps * sizeof(double)), "hipMalloc");
    struct Free {
        double *p;
        ~Free() { (void)hipFree(p); }
    } release{d_out};
    check(hipMemcpyAsync(h->rec_signal, signal, steps * sizeof(float), hipMemcpyHostToDevice, h->stream), "hipMemcpyAsync");
    const int code = as_fdtd_reset(h);
    if (code != AS_OK) throw Failure{code, h->error};
    // Keep the history where asked, or by default where it fits the budget.
    const size_t G = std::max<size_t>(h->specials, 1);
    const size_t hist_bytes = steps * G * e;
    bool keep = h->keep_history > 0;
    if (h->keep_history == 0) {
        size_t free_bytes = 0, total_bytes = 0;
        check(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
        keep = hist_bytes <= std::min(kHistoryBudget, free_bytes / 4);
    }
    if (h->layer_cells > 0 && !keep) {
        // Its checkpoints hold no layer state to recompute a segment from.
        refuse("the adjoint of a grid with a matched layer keeps the record's history: keep_history off, or a "
               "history beyond 16 GiB or a quarter of free device memory, is refused");
    }
    StepArgs<T> hs = step_args<T>(h);
    const uint32_t hblocks = block_count(h, hs);
    if (keep) {
        check(hipMalloc(&h->rec_hist, hist_bytes), "hipMalloc");
        if (h->rec_graph_every != every) {
            free_record_graphs(h);
            check(hipMalloc(reinterpret_cast<void **>(&h->rec_sig_stage), every * sizeof(float)), "hipMalloc");
            check(hipMalloc(reinterpret_cast<void **>(&h->rec_out_stage), std::max<size_t>(P, 1) * every * sizeof(double)),
                  "hipMalloc");
            check(hipMalloc(&h->rec_hist_stage, every * G * e), "hipMalloc");
            h->rec_graph_every = every;
        }
    }
    for (size_t seg = 0; seg < segments; ++seg) {
        char *base = h->rec_ckpt + seg * per;
        // By role: p(k - 1), p(k), w(k - 1), w(k).
        check(hipMemcpyAsync(base, h->field[h->prev], field, hipMemcpyDeviceToDevice, h->stream), "ckpt");
        check(hipMemcpyAsync(base + field, h->field[h->prev ^ 1], field, hipMemcpyDeviceToDevice, h->stream), "ckpt");
        if (states) {
            check(hipMemcpyAsync(base + 2 * field, h->velocity[h->prev], states, hipMemcpyDeviceToDevice, h->stream), "ckpt");
            check(hipMemcpyAsync(base + 2 * field + states, h->velocity[h->prev ^ 1], states, hipMemcpyDeviceToDevice,
                                 h->stream),
                  "ckpt");
        }
        const size_t k0 = seg * every, k1 = std::min(steps, k0 + every);
        if (!keep) {
            launch_steps<T>(h, h->rec_signal + k0, k1 - k0, d_out + k0, steps);
            continue;
        }
        T *hist = static_cast<T *>(h->rec_hist) + k0 * G;
        if (k1 - k0 < every || hblocks == 0) {
            StepArgs<T> d = hs;
            d.signal = h->rec_signal + k0;
            d.out = d_out + k0;
            d.out_stride = steps;
            d.history = hist;
            if (hblocks) h->prev = record_steps<T>(h, d, hblocks, h->prev, k1 - k0);
            h->done += k1 - k0;
            continue;
        }
        // A whole segment: its graph (by starting slot) on the staging buffers.
        const int slot = h->prev;
        if (h->rec_graph[slot] == nullptr) {
            StepArgs<T> g = hs;
            g.signal = h->rec_sig_stage;
            g.out = h->rec_out_stage;
            g.out_stride = every;
            g.history = static_cast<T *>(h->rec_hist_stage);
            capture_graph(h->stream, &h->rec_graph[slot], [&] { (void)record_steps<T>(h, g, hblocks, slot, every); });
        }
        check(hipMemcpyAsync(h->rec_sig_stage, h->rec_signal + k0, every * sizeof(float), hipMemcpyDeviceToDevice,
                             h->stream),
              "hipMemcpyAsync");
        check(hipGraphLaunch(h->rec_graph[slot], h->stream), "hipGraphLaunch");
        if (P > 0) {
            check(hipMemcpy2DAsync(d_out + k0, steps * sizeof(double), h->rec_out_stage, every * sizeof(double),
                                   every * sizeof(double), P, hipMemcpyDeviceToDevice, h->stream),
                  "hipMemcpy2DAsync");
        }
        check(hipMemcpyAsync(hist, h->rec_hist_stage, every * G * e, hipMemcpyDeviceToDevice, h->stream), "history");
        h->prev = static_cast<int>(slot ^ (every & 1u));
        h->done += every;
    }
    if (out != nullptr && P > 0) {
        check(hipMemcpyAsync(out, d_out, P * steps * sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
    }
    check(hipStreamSynchronize(h->stream), "record");
    h->rec_steps = steps;
    h->rec_every = every;
}

// The backward half: segment by segment from the last, the segment recomputed from its
// checkpoint (keeping the specials' p(n + 1) - p(n - 1)), then the adjoint stepped back
// through it; the sums of the gradient, into sens_g and sens_gain.
template <typename T>
void adjoint(as_fdtd *h, const double *dl_dout, double *sens_g, double *sens_gain) {
    check_adjoint(h);
    if (h->rec_ckpt == nullptr) refuse("no forward run recorded (as_fdtd_record)");
    if (h->layer_cells > 0 && h->rec_hist == nullptr) refuse("the adjoint of a grid with a matched layer needs the record's history");
    const size_t steps = h->rec_steps, every = h->rec_every;
    const size_t e = sizeof(T);
    const size_t G = h->specials, slots = h->state_slots, P = h->probes;
    const size_t segments = (steps + every - 1) / every;
    std::vector<void *> temporaries;
    auto alloc = [&](size_t bytes) {
        void *p = nullptr;
        check(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc");
        temporaries.push_back(p);
        check(hipMemsetAsync(p, 0, std::max<size_t>(bytes, 256), h->stream), "hipMemsetAsync");
        return p;
    };
    struct Free {
        std::vector<void *> &v;
        ~Free() {
            for (void *p : v) (void)hipFree(p);
        }
    } release{temporaries};
    const size_t field = h->n_pad * e, states = slots * e, per = 2 * field + 2 * states;
    T *history = static_cast<T *>(alloc(every * std::max<size_t>(G, 1) * e));
    T *q0 = static_cast<T *>(alloc(field)), *q1 = static_cast<T *>(alloc(field));
    T *v0 = static_cast<T *>(alloc(states)), *v1 = static_cast<T *>(alloc(states));
    double *acc_g = static_cast<double *>(alloc(std::max<size_t>(G, 1) * sizeof(double)));
    double *acc_gain = static_cast<double *>(alloc(std::max<size_t>(slots, 1) * sizeof(double)));
    double *d_g = static_cast<double *>(alloc(std::max<size_t>(P, 1) * steps * sizeof(double)));
    if (P > 0) {
        check(hipMemcpyAsync(d_g, dl_dout, P * steps * sizeof(double), hipMemcpyHostToDevice, h->stream),
              "hipMemcpyAsync");
    }
    // The probes' taps, by distinct cell, each with its coef_inv (1 off the special cells).
    std::vector<std::pair<uint32_t, std::pair<uint32_t, double>>> taps;  // (cell, (probe, weight))
    for (uint32_t p = 0; p < P; ++p)
        for (uint32_t t = h->probe_start_host[p]; t < h->probe_start_host[p + 1]; ++t)
            taps.push_back({h->probe_cell_host[t], {p, h->probe_weight_host[t]}});
    std::stable_sort(taps.begin(), taps.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<uint32_t> tcell, tstart{0}, tprobe;
    std::vector<double> tweight;
    std::vector<double> tci;
    std::vector<std::pair<uint64_t, double>> ci_of;
    for (size_t k = 0; k < h->special_cell_host.size(); ++k) ci_of.push_back({h->special_cell_host[k], h->coef_inv_host[k]});
    std::sort(ci_of.begin(), ci_of.end());
    for (size_t t = 0; t < taps.size(); ++t) {
        if (t == 0 || taps[t].first != taps[t - 1].first) {
            if (t) tstart.push_back(static_cast<uint32_t>(tprobe.size()));
            tcell.push_back(taps[t].first);
            const auto it = std::lower_bound(ci_of.begin(), ci_of.end(), std::make_pair<uint64_t, double>(taps[t].first, -1.0));
            tci.push_back(it != ci_of.end() && it->first == taps[t].first ? it->second : 1.0);
        }
        tprobe.push_back(taps[t].second.first);
        tweight.push_back(taps[t].second.second);
    }
    tstart.push_back(static_cast<uint32_t>(tprobe.size()));
    const uint32_t unique = static_cast<uint32_t>(tcell.size());
    std::vector<float> scale_host;
    if (h->scaled) {
        scale_host.resize(h->n_pad);
        check(hipMemcpy(scale_host.data(), h->scale, h->n_pad * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    // The layer's cells by padded cell (the layer's adjoint keeps its own variables).
    const size_t layer_count = h->layer_cells;
    std::vector<uint32_t> layer_index;
    if (layer_count > 0) {
        layer_index.assign(h->n_pad, kNoTap);
        for (uint32_t c = 0; c < layer_count; ++c) layer_index[h->layer_cell_host[c]] = c;
    }
    for (uint32_t u = 0; u < unique; ++u) {
        if (!layer_index.empty() && layer_index[tcell[u]] != kNoTap) {
            tci[u] = 1.0;  // a layer cell's own adjoint: dJ/dy as it is
        } else if (h->scaled) {
            tci[u] *= static_cast<double>(scale_host[tcell[u]]);  // R = s Q: times the cell's scale
        }
    }
    auto put = [&](const auto &v) {
        using U = typename std::decay_t<decltype(v)>::value_type;
        U *p = static_cast<U *>(alloc(std::max<size_t>(v.size(), 1) * sizeof(U)));
        if (!v.empty()) check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice, h->stream), "hipMemcpyAsync");
        return p;
    };
    uint32_t *d_tstart = put(tstart), *d_tprobe = put(tprobe);
    double *d_tweight = put(tweight);
    double *d_tci = put(tci);
    // Where each tap cell's dJ/dy enters: a special adds it in its own update (sp_tap); an
    // interior cell becomes kClassTap in the adjoint's own copy of the classes, and its tile
    // is no longer taken as all interior, so the column march adds it (tap_of). Solid cells
    // are read by no update: their taps are left out.
    std::vector<uint8_t> types_host(h->n_pad);
    check(hipMemcpy(types_host.data(), h->types, h->n_pad, hipMemcpyDeviceToHost), "hipMemcpy");
    std::vector<uint32_t> tiles_host(h->tile_count);
    if (h->tile_count) {
        check(hipMemcpy(tiles_host.data(), h->tiles, h->tile_count * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    std::vector<std::pair<uint64_t, uint32_t>> special_at;  // (cell, sorted special)
    for (size_t k = 0; k < G; ++k) special_at.push_back({h->special_cell_host[h->sorted_to_special[k]], static_cast<uint32_t>(k)});
    std::sort(special_at.begin(), special_at.end());
    std::vector<uint32_t> sp_tap(G, kNoTap), tap_tiles, layer_tap;
    if (layer_count > 0) layer_tap.assign(layer_count, kNoTap);
    std::vector<std::pair<uint32_t, uint32_t>> interior_taps;  // (cell, tap)
    for (uint32_t u = 0; u < unique; ++u) {
        const uint64_t c = tcell[u];
        const auto it = std::lower_bound(special_at.begin(), special_at.end(), std::make_pair(c, uint32_t{0}));
        if (it != special_at.end() && it->first == c) {
            sp_tap[it->second] = u;
        } else if (!layer_index.empty() && layer_index[c] != kNoTap) {
            layer_tap[layer_index[c]] = u;
        } else if (types_host[c] == kClassInterior) {
            types_host[c] = kClassTap;
            interior_taps.push_back({static_cast<uint32_t>(c), u});
            const uint64_t x = c % static_cast<uint64_t>(h->px), y = (c / static_cast<uint64_t>(h->px)) % static_cast<uint64_t>(h->ny),
                           z = c / static_cast<uint64_t>(h->sxy);
            tap_tiles.push_back(static_cast<uint32_t>(x / kTileX + h->tiles_x * ((y - 1) / kTileRows + h->tiles_y * ((z - 1) / h->tile_z))));
        }
    }
    std::sort(tap_tiles.begin(), tap_tiles.end());
    for (auto &entry : tiles_host) {
        if (entry != kNoTile && std::binary_search(tap_tiles.begin(), tap_tiles.end(), entry & ~kAllInterior)) entry &= ~kAllInterior;
    }
    uint8_t *d_types = nullptr;
    uint32_t *d_tiles = nullptr, *d_tap_of = nullptr;
    if (!interior_taps.empty()) {
        d_types = put(types_host);
        d_tiles = put(tiles_host);
        d_tap_of = static_cast<uint32_t *>(alloc(h->n_pad * sizeof(uint32_t)));
        std::vector<uint32_t> tap_of(h->n_pad, 0);  // read at the tap cells only
        for (const auto &ct : interior_taps) tap_of[ct.first] = ct.second;
        check(hipMemcpyAsync(d_tap_of, tap_of.data(), h->n_pad * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream),
              "hipMemcpyAsync");
        check(hipStreamSynchronize(h->stream), "adjoint");  // before tap_of goes
    }
    uint32_t *d_sp_tap = G ? put(sp_tap) : nullptr;
    // The matched layer's adjoint (see LayerAdjoint): its couplings with the cells off the
    // layer next to it, each with the coefficient that cell's row puts on the layer cell's
    // p (in the fields' variables: Q, or R = s Q) and the factor that takes the layer's term
    // into that cell's field.
    LayerAdjoint<T> la{};
    if (layer_count > 0) {
        std::vector<std::pair<uint64_t, uint32_t>> special_of;  // (cell, layout special)
        for (size_t k = 0; k < h->special_cell_host.size(); ++k) special_of.push_back({h->special_cell_host[k], static_cast<uint32_t>(k)});
        std::sort(special_of.begin(), special_of.end());
        std::vector<uint32_t> cstart{0}, ccell;
        std::vector<uint8_t> cdir;
        std::vector<T> calpha, cmult;
        const long long stride[3] = {1, h->px, static_cast<long long>(h->sxy)};
        for (uint32_t c = 0; c < layer_count; ++c) {
            const uint64_t i = h->layer_cell_host[c];
            for (uint32_t dir = 0; dir < 6; ++dir) {
                const long long step = (dir & 1u) ? stride[dir / 2] : -stride[dir / 2];
                const uint64_t j = static_cast<uint64_t>(static_cast<long long>(i) + step);
                if (layer_index[j] != kNoTap || types_host[j] == 0) continue;  // a layer cell, or solid
                const double s_j = h->scaled ? static_cast<double>(scale_host[j]) : 1.0;
                double alpha = 0.0, mult = 0.0;
                const auto it = std::lower_bound(special_of.begin(), special_of.end(), std::make_pair(j, uint32_t{0}));
                if (it != special_of.end() && it->first == j) {
                    // Its row: coef_inv (... + w p_layer ...): w on Q, w / s on R; its field
                    // takes coef_inv (times s) of the layer's term.
                    const uint32_t k = it->second;
                    alpha = static_cast<double>(h->special_axial_host[6 * static_cast<size_t>(k) + (dir ^ 1u)]) / s_j;
                    mult = h->coef_inv_host[k] * s_j;
                } else {
                    // An interior row: s axial on p_layer, axial on R (axial on Q unscaled).
                    alpha = h->axial;
                    mult = s_j;
                }
                cdir.push_back(static_cast<uint8_t>(dir));
                ccell.push_back(static_cast<uint32_t>(j));
                calpha.push_back(static_cast<T>(alpha));
                cmult.push_back(static_cast<T>(mult));
            }
            cstart.push_back(static_cast<uint32_t>(ccell.size()));
        }
        la.cells = static_cast<uint32_t>(layer_count);
        la.cell = h->layer_cell;
        la.index = put(layer_index);
        for (int k = 0; k < 6; ++k) {
            la.keep[k] = h->keep[k];
            la.push[k] = h->push[k];
        }
        la.scale = h->layer_scale;
        la.px = static_cast<int>(h->px);
        la.ny = static_cast<int>(h->ny);
        la.sxy = static_cast<long long>(h->sxy);
        la.part = static_cast<T *>(alloc(3 * layer_count * e));
        la.face = static_cast<T *>(alloc(6 * layer_count * e));
        la.couple_start = put(cstart);
        la.couple_dir = put(cdir);
        la.couple_cell = put(ccell);
        la.couple_alpha = put(calpha);
        la.couple_mult = put(cmult);
        la.tap = put(layer_tap);
    }
    const uint32_t layer_blocks = static_cast<uint32_t>((layer_count + kBlock - 1) / kBlock);
    T *inj = static_cast<T *>(alloc((every + 1) * std::max<uint32_t>(unique, 1) * e));

    StepArgs<T> r = step_args<T>(h);  // recompute: no probes, history on
    r.probes = 0;
    r.probe_blocks = 0;
    r.history = history;
    const uint32_t rblocks = block_count(h, r);
    StepArgs<T> a = step_args<T>(h);  // adjoint
    a.probes = 0;
    a.probe_blocks = 0;
    a.acc_g = acc_g;
    a.acc_gain = acc_gain;
    if (d_types != nullptr) {
        a.types = d_types;
        a.tiles = d_tiles;
        a.tap_of = d_tap_of;
    }
    a.sp_tap = d_sp_tap;
    const uint32_t ablocks = a.special_blocks + a.tile_blocks;
    const uint32_t sblocks = (h->specials + kBlock - 1) / kBlock;

    // Segment k0 .. k1 - 1 recomputed from slots 0 (p(k0 - 1)) and 1 (p(k0)), its drive at sig.
    auto recompute = [&](size_t count, const float *sig) {
        r.signal = sig;
        for (size_t k = 0; k < count; ++k) launch_one<T>(h, r, rblocks, static_cast<int>(k & 1u), k, h->stream);
    };
    // Adjoint steps n = k0 + top down to k0 + 1, each adding the previous step's sums
    // (none before the first), dJ/dy of step n at g[probe * gstride + n - gbase]; then
    // the last step's sums. m counts the adjoint steps before: its parity says which
    // field is which.
    auto backward = [&](size_t k0, size_t top, const double *g, size_t gstride, size_t gbase, unsigned long long m) {
        // The segment's dJ/dy through the taps' weights, for n = k0 + 1 .. k0 + top (row
        // n - k0 - 1 of inj).
        if (unique > 0 && top > 0) {
            const uint32_t rows = static_cast<uint32_t>(top), count = unique * rows;
            hipLaunchKernelGGL((tap_values<T>), dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, h->stream, inj, d_tci,
                               d_tstart, d_tprobe, d_tweight, g, static_cast<unsigned long long>(gstride),
                               static_cast<unsigned long long>(k0 + 1 - gbase), unique, rows);
        }
        for (size_t n = k0 + top; n >= k0 + 1; --n, ++m) {
            const bool even = (m & 1ull) == 0;
            a.a = even ? q0 : q1;
            a.o = even ? q0 : q1;  // in place
            a.b = even ? q1 : q0;
            a.wa = even ? v0 : v1;
            a.wo = even ? v0 : v1;
            a.wb = even ? v1 : v0;
            a.delta = n < k0 + top ? history + (n - k0) * G : nullptr;
            a.inj = inj + (n - k0 - 1) * unique;
            if (ablocks) {
                if (h->wide && h->scaled) {
                    hipLaunchKernelGGL((adjoint_kernel<T, true, true>), dim3(ablocks), dim3(kBlock), 0, h->stream, a);
                } else if (h->wide) {
                    hipLaunchKernelGGL((adjoint_kernel<T, true, false>), dim3(ablocks), dim3(kBlock), 0, h->stream, a);
                } else if (h->scaled) {
                    hipLaunchKernelGGL((adjoint_kernel<T, false, true>), dim3(ablocks), dim3(kBlock), 0, h->stream, a);
                } else {
                    hipLaunchKernelGGL((adjoint_kernel<T, false, false>), dim3(ablocks), dim3(kBlock), 0, h->stream, a);
                }
            }
            if (layer_blocks) {
                la.b = a.b;
                la.o = a.o;
                la.inj = a.inj;
                hipLaunchKernelGGL((layer_pressure_adjoint<T>), dim3(layer_blocks), dim3(kBlock), 0, h->stream, la);
                hipLaunchKernelGGL((layer_faces_adjoint<T>), dim3(layer_blocks), dim3(kBlock), 0, h->stream, la);
            }
        }
        if (sblocks) {
            if (h->scaled) {
                hipLaunchKernelGGL((accumulate_kernel<T, true>), dim3(sblocks), dim3(kBlock), 0, h->stream, a, a.a, a.wa,
                                   history, acc_g, acc_gain);
            } else {
                hipLaunchKernelGGL((accumulate_kernel<T, false>), dim3(sblocks), dim3(kBlock), 0, h->stream, a, a.a,
                                   a.wa, history, acc_g, acc_gain);
            }
        }
    };
    // Whole segments (every steps, the adjoint's top at the segment's end) replay captured
    // graphs: the recompute from a staged drive, the adjoint from staged dJ/dy, one graph
    // for each parity of the adjoint steps before.
    float *sig_stage = static_cast<float *>(alloc(every * sizeof(float)));
    double *g_stage = static_cast<double *>(alloc(std::max<size_t>(P, 1) * (every + 1) * sizeof(double)));
    std::vector<hipGraphExec_t> graphs(3, nullptr);  // recompute, adjoint from an even and an odd m
    struct Graphs {
        std::vector<hipGraphExec_t> &v;
        ~Graphs() {
            for (auto g : v)
                if (g) (void)hipGraphExecDestroy(g);
        }
    } release_graphs{graphs};
    auto capture = [&](int which, auto &&enqueue) {
        if (graphs[which] != nullptr) return;
        capture_graph(h->stream, &graphs[which], enqueue);
    };
    unsigned long long m = 0;
    for (size_t seg = segments; seg-- > 0;) {
        const size_t k0 = seg * every, k1 = std::min(steps, k0 + every);
        // The segment's history: the record's, copied in where it kept it; else recomputed
        // from the checkpoint, restored into slots 0 (p(k0 - 1)) and 1 (p(k0)).
        const bool kept = h->rec_hist != nullptr;
        if (kept) {
            const size_t rows = std::max<size_t>(G, 1) * e;
            check(hipMemcpyAsync(history, static_cast<const char *>(h->rec_hist) + k0 * rows, (k1 - k0) * rows,
                                 hipMemcpyDeviceToDevice, h->stream),
                  "history");
        } else {
            const char *base = h->rec_ckpt + seg * per;
            check(hipMemcpyAsync(h->field[0], base, field, hipMemcpyDeviceToDevice, h->stream), "ckpt");
            check(hipMemcpyAsync(h->field[1], base + field, field, hipMemcpyDeviceToDevice, h->stream), "ckpt");
            if (states) {
                check(hipMemcpyAsync(h->velocity[0], base + 2 * field, states, hipMemcpyDeviceToDevice, h->stream), "ckpt");
                check(hipMemcpyAsync(h->velocity[1], base + 2 * field + states, states, hipMemcpyDeviceToDevice,
                                     h->stream),
                      "ckpt");
            }
        }
        // Adjoint steps n from min(k1, steps - 1) down to k0 + 1 (Q(steps) is 0: p(steps) is
        // read by no probe; Q(0) enters no gradient).
        const size_t top = std::min(k1, steps - 1);
        const bool whole = k1 - k0 == every && top == k1 && h->graph_steps > 0;
        if (!whole) {
            if (!kept) recompute(k1 - k0, h->rec_signal + k0);
            if (top > k0) backward(k0, top - k0, d_g, steps, 0, m);
            m += top > k0 ? top - k0 : 0;
            continue;
        }
        check(hipMemcpyAsync(sig_stage, h->rec_signal + k0, every * sizeof(float), hipMemcpyDeviceToDevice, h->stream),
              "hipMemcpyAsync");
        if (!kept) {
            capture(0, [&] { recompute(every, sig_stage); });
            check(hipGraphLaunch(graphs[0], h->stream), "hipGraphLaunch");
        }
        if (P > 0) {
            check(hipMemcpy2DAsync(g_stage, (every + 1) * sizeof(double), d_g + k0, steps * sizeof(double),
                                   (every + 1) * sizeof(double), P, hipMemcpyDeviceToDevice, h->stream),
                  "hipMemcpy2DAsync");
        }
        const int which = 1 + static_cast<int>(m & 1ull);
        const unsigned long long start = m;
        capture(which, [&] { backward(0, every, g_stage, every + 1, 0, start); });
        check(hipGraphLaunch(graphs[which], h->stream), "hipGraphLaunch");
        m += every;
    }
    check(hipGetLastError(), "adjoint");
    std::vector<double> g(G), gain(slots);
    if (G) check(hipMemcpyAsync(g.data(), acc_g, G * sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
    if (slots) check(hipMemcpyAsync(gain.data(), acc_gain, slots * sizeof(double), hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(h->stream), "adjoint");
    if (sens_g != nullptr) {
        std::fill(sens_g, sens_g + h->special_cell_host.size(), 0.0);
        for (size_t k = 0; k < G; ++k) sens_g[h->sorted_to_special[k]] = -g[k];
    }
    if (sens_gain != nullptr) {
        for (size_t j = 0; j < h->state_slot.size(); ++j) sens_gain[j] = h->state_slot[j] == ~0ull ? 0.0 : gain[h->state_slot[j]];
    }
    free_record(h);
}

}  // namespace

extern "C" {

int as_fdtd_record(as_fdtd *h, const float *signal, size_t steps, size_t every, double *out) {
    if (h == nullptr) return AS_EINVAL;
    try {
        if (steps < 2) refuse("the adjoint needs at least two steps");
        if (signal == nullptr) refuse("the record needs a drive");
        if (h->precision == 64) {
            record<double>(h, signal, steps, every, out);
        } else {
            record<float>(h, signal, steps, every, out);
        }
        return AS_OK;
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_adjoint(as_fdtd *h, const double *dl_dout, double *sens_g, double *sens_gain) {
    if (h == nullptr) return AS_EINVAL;
    try {
        if (dl_dout == nullptr && h->probes > 0) refuse("the adjoint needs dJ/dy");
        if (h->precision == 64) {
            adjoint<double>(h, dl_dout, sens_g, sens_gain);
        } else {
            adjoint<float>(h, dl_dout, sens_g, sens_gain);
        }
        // Leave the handle at rest, as after reset().
        return as_fdtd_reset(h);
    } catch (const Failure &f) {
        return fail(h, f);
    }
}

int as_fdtd_gradient(as_fdtd *h, const float *signal, size_t steps, size_t every, const double *dl_dout, double *out,
                     double *sens_g, double *sens_gain) {
    if (h == nullptr) return AS_EINVAL;
    if (dl_dout == nullptr && h->probes > 0) {
        h->error = "the gradient needs a drive and dJ/dy";
        return AS_EINVAL;
    }
    const int code = as_fdtd_record(h, signal, steps, every, out);
    if (code != AS_OK) return code;
    return as_fdtd_adjoint(h, dl_dout, sens_g, sens_gain);
}

}  // extern "C"

extern "C" {

void *as_fdtd_stream(as_fdtd *h) { return h ? static_cast<void *>(h->stream) : nullptr; }

int as_fdtd_info(const as_fdtd *h, uint64_t *out) {
    if (h == nullptr || out == nullptr) return AS_EINVAL;
    out[0] = h->interior_cells;
    out[1] = static_cast<uint64_t>(h->specials) + h->face_specials;
    out[2] = h->states;
    out[3] = h->layer_cells;
    out[4] = h->probes;
    out[5] = h->tile_count;
    out[6] = static_cast<uint64_t>(h->px);
    out[7] = h->done;
    out[8] = h->pairs ? static_cast<uint64_t>(h->pair_z) : 0u;
    out[9] = h->pairs ? h->cell_count : 0u;
    out[10] = h->sp_rec != nullptr ? h->kinds : 0u;
    return AS_OK;
}

const char *as_fdtd_last_error(const as_fdtd *h) { return h ? h->error.c_str() : "no handle"; }

void as_fdtd_destroy(as_fdtd *h) {
    if (h == nullptr) return;
    // Work in flight ends first; then the graphs go, before any buffer their kernel nodes
    // point at is freed; the stream last.
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &g : h->graph_exec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    free_record_graphs(h);
    free_record(h);
    for (void *p : h->allocations) (void)hipFree(p);
    if (h->signal_buf) (void)hipFree(h->signal_buf);
    if (h->out_buf) (void)hipFree(h->out_buf);
    if (h->sig_stage) (void)hipFree(h->sig_stage);
    if (h->out_stage) (void)hipFree(h->out_stage);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
