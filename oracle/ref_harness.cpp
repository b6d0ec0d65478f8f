// Test infrastructure only: a C-ABI wrapper round the reference's own header-only wave
// solver (acoustic_spaces::WaveSolver, cpp/include/acoustic_spaces/fdtd.hpp:339-3601),
// compiled from the reference's headers where they lie (oracle/Makefile) into
// oracle/_ref/libasref.so. Tests use it to pin the oracle and the product's layout
// producer; bench.py's cpu_baseline leg times its run(). Nothing in the product links it.
//
// The constructor's arguments follow the nanobind binding make_wave_solver
// (cpp/bindings/native_module.cpp:1238-1285); the layout's arrays are captured once, by
// WaveSolver::layout (fdtd.hpp:727-782), as the binding's wave_layout does (1297-1304).

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "acoustic_spaces/fdtd.hpp"

namespace {

namespace as = acoustic_spaces;

// Element kinds of a layout entry, as the Python side reads them.
enum Kind : int { kU8 = 0, kI32 = 1, kI64 = 2, kU64 = 3, kF32 = 4, kF64 = 5 };

template <typename T>
constexpr int kind_of() {
    if constexpr (std::is_same_v<T, std::uint8_t>) return kU8;
    if constexpr (std::is_same_v<T, std::int32_t>) return kI32;
    if constexpr (std::is_same_v<T, std::int64_t>) return kI64;
    if constexpr (std::is_same_v<T, std::size_t>) return kU64;
    if constexpr (std::is_same_v<T, float>) return kF32;
    if constexpr (std::is_same_v<T, double>) return kF64;
    return -1;
}

struct Entry {
    std::string name;
    int kind = -1;
    std::size_t count = 0;
    std::vector<unsigned char> bytes;
};

struct Handle {
    std::unique_ptr<as::Bvh> bvh;
    std::unique_ptr<as::WaveSolver> solver;
    std::vector<Entry> layout;
};

void put_error(char *err, std::size_t len, const char *what) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", what);
    }
}

}  // namespace

extern "C" {

void *asref_create(const double *vertices, std::size_t nv, const std::int64_t *faces, std::size_t nf,
                   const std::int64_t *face_material, const double *beta, std::size_t materials,
                   const double *branches, const std::int64_t *branch_material, std::size_t nbranch,
                   const double *origin, double cell, std::size_t nx, std::size_t ny, std::size_t nz,
                   const double *source, const double *probes, std::size_t nprobes,
                   double edge_admittance, std::size_t layer, double layer_damping, std::size_t threads,
                   std::size_t stencil, const float *speed, const float *flow_x, const float *flow_y,
                   std::size_t probes_per_listener, int fitted, char *err, std::size_t errlen) {
    try {
        auto h = std::make_unique<Handle>();
        h->bvh = std::make_unique<as::Bvh>(vertices, nv, faces, nf);
        std::vector<as::Vec3> points;
        for (std::size_t p = 0; p < nprobes; ++p) {
            points.push_back(as::Vec3{probes[3 * p], probes[3 * p + 1], probes[3 * p + 2]});
        }
        as::CutCells::Mesh mesh;
        mesh.vertices = vertices;
        mesh.vertex_count = nv;
        mesh.faces = faces;
        mesh.face_count = nf;
        h->solver = std::make_unique<as::WaveSolver>(
            *h->bvh, face_material, beta, materials, branches, branch_material, nbranch,
            as::Vec3{origin[0], origin[1], origin[2]}, cell, nx, ny, nz,
            as::Vec3{source[0], source[1], source[2]}, points, edge_admittance, layer, layer_damping,
            threads, stencil, speed, flow_x, flow_y, probes_per_listener, fitted != 0,
            fitted != 0 ? &mesh : nullptr);
        h->solver->layout([&](const char *name, const auto &values) {
            using T = typename std::decay_t<decltype(values)>::value_type;
            Entry e;
            e.name = name;
            e.kind = kind_of<T>();
            e.count = values.size();
            e.bytes.resize(values.size() * sizeof(T));
            if (!values.empty()) std::memcpy(e.bytes.data(), values.data(), e.bytes.size());
            h->layout.push_back(std::move(e));
        });
        return h.release();
    } catch (const std::exception &e) {
        put_error(err, errlen, e.what());
        return nullptr;
    }
}

int asref_layout_count(void *handle) { return static_cast<int>(static_cast<Handle *>(handle)->layout.size()); }

int asref_layout_entry(void *handle, int k, const char **name, int *kind, std::size_t *count,
                       const void **data) {
    auto *h = static_cast<Handle *>(handle);
    if (k < 0 || static_cast<std::size_t>(k) >= h->layout.size()) return -1;
    const Entry &e = h->layout[static_cast<std::size_t>(k)];
    *name = e.name.c_str();
    *kind = e.kind;
    *count = e.count;
    *data = e.bytes.data();
    return 0;
}

void asref_run(void *handle, const float *signal, std::size_t steps, double *out) {
    static_cast<Handle *>(handle)->solver->run(signal, steps, out);
}

void asref_reset(void *handle) { static_cast<Handle *>(handle)->solver->reset(); }

void asref_destroy(void *handle) { delete static_cast<Handle *>(handle); }

// air cells, boundary cells, branch states, layer cells, probes, threads, open, source on face
void asref_info(void *handle, std::size_t *out) {
    const as::WaveSolver &s = *static_cast<Handle *>(handle)->solver;
    out[0] = s.air_cells();
    out[1] = s.boundary_cells();
    out[2] = s.branch_states();
    out[3] = s.layer_cells();
    out[4] = s.probes();
    out[5] = s.threads();
    out[6] = s.open() ? 1 : 0;
    out[7] = s.source_on_face() ? 1 : 0;
}

// moving air: flowing, wind_round, wind_fastest, wind_sweeps, mean_square (of the last run)
void asref_flow(void *handle, double *out) {
    const as::WaveSolver &s = *static_cast<Handle *>(handle)->solver;
    out[0] = s.flowing() ? 1.0 : 0.0;
    out[1] = s.wind_round() ? 1.0 : 0.0;
    out[2] = s.wind_fastest();
    out[3] = static_cast<double>(s.wind_sweeps());
    out[4] = s.mean_square();
}

// Fitted walls: the fit's rows (CutCells::Results, cut_cells.hpp:1249-1265) as the solver
// keeps them (fdtd.hpp:723). Returns the row count (-1 without fitted walls); where the
// pointers are not null, copies the rows' cells (n), air volumes (n), open areas (18 n), wall
// areas (n (materials + 1)) and slivers (6 n), and a few of the fit's counts into stats:
// cubes, regions, ties, farthest, conduit links, conduit scaled, far volume, diagonal area,
// dropped area, edge area, air volume, least conduit factor, conduit wall moved.
long asref_cut_rows(void *handle, std::int64_t *cells, double *volume, double *open, double *wall, double *sliver,
                    double *stats) {
    const as::CutCells::Results *cut = static_cast<Handle *>(handle)->solver->cut_cells();
    if (cut == nullptr) return -1;
    const std::size_t n = cut->cells().size();
    auto put = [](auto *dst, const auto &src) {
        if (dst != nullptr && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    put(cells, cut->cells());
    put(volume, cut->air_volume());
    put(open, cut->open_area());
    put(wall, cut->wall_area());
    put(sliver, cut->sliver());
    if (stats != nullptr) {
        const auto &s = cut->stats();
        const double v[] = {static_cast<double>(s.cubes), static_cast<double>(s.regions), static_cast<double>(s.ties),
                            static_cast<double>(s.farthest), static_cast<double>(s.conduit_links),
                            static_cast<double>(s.conduit_scaled), s.far_volume, s.diagonal_area, s.dropped_area,
                            s.edge_area, s.air_volume, s.conduit_least, s.conduit_wall_moved};
        std::memcpy(stats, v, sizeof v);
    }
    return static_cast<long>(n);
}

// Fitted walls: the solver's counts of the sliver rule and the clamp (fdtd.hpp:1294-1301):
// raised, most raised, sliver links, both, thick, layer whole links, plain rows.
void asref_fitted_counts(void *handle, double *out) {
    const auto &c = static_cast<Handle *>(handle)->solver->fitted_counts();
    const double v[] = {static_cast<double>(c.raised), c.most_raised, static_cast<double>(c.sliver_links),
                        static_cast<double>(c.sliver_both), static_cast<double>(c.sliver_thick),
                        static_cast<double>(c.layer_whole_links), static_cast<double>(c.plain_rows)};
    std::memcpy(out, v, sizeof v);
}

void asref_open_sides(void *handle, int *out) {
    const auto &sides = static_cast<Handle *>(handle)->solver->open_sides();
    for (int k = 0; k < 6; ++k) out[k] = sides[static_cast<std::size_t>(k)] ? 1 : 0;
}

int asref_probe_found(void *handle, std::size_t p) {
    return static_cast<Handle *>(handle)->solver->probe_found(p) ? 1 : 0;
}

int asref_probe_in_sight(void *handle, std::size_t p) {
    return static_cast<Handle *>(handle)->solver->probe_in_sight(p) ? 1 : 0;
}

}  // extern "C"
