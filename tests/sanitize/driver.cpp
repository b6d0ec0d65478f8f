// Test infrastructure: runs the host set-up (csrc/setup.cpp, csrc/fitted.cpp, through the C
// ABI) on a scene read from a file, in a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitized_setup.py writes the scenes and builds
// this). Prints the layout's counts, or the refusal; exits 0 either way, so that only a
// sanitizer's report (which aborts) fails the test.
//
// The scene file: little-endian, each array preceded by its element count (uint64):
// vertices (f64 x 3), faces (i64 x 3), face_material (i64), beta (f64), branches (f64 x 3),
// branch_material (i64), origin (f64 x 3), source (f64 x 3), probes (f64 x 3), speed (f32),
// flow_x (f32), flow_y (f32); then cell (f64), nx ny nz layer stencil per_listener flags
// (uint64 each), layer_damping (f64).
#include <cstdint>
#include <cstdio>
#include <vector>

#include "asfdtd.h"

template <typename T>
static std::vector<T> read_array(FILE *f, size_t width) {
    uint64_t count = 0;
    if (std::fread(&count, sizeof count, 1, f) != 1) return {};
    std::vector<T> v(count * width);
    if (!v.empty() && std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) v.clear();
    return v;
}

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    const auto vertices = read_array<double>(f, 3);
    const auto faces = read_array<int64_t>(f, 3);
    const auto face_material = read_array<int64_t>(f, 1);
    const auto beta = read_array<double>(f, 1);
    const auto branches = read_array<double>(f, 3);
    const auto branch_material = read_array<int64_t>(f, 1);
    const auto origin = read_array<double>(f, 3);
    const auto source = read_array<double>(f, 3);
    const auto probes = read_array<double>(f, 3);
    const auto speed = read_array<float>(f, 1);
    const auto flow_x = read_array<float>(f, 1);
    const auto flow_y = read_array<float>(f, 1);
    double cell = 0, damping = 0;
    uint64_t u[7] = {};
    if (std::fread(&cell, sizeof cell, 1, f) != 1 || std::fread(u, sizeof u[0], 7, f) != 7 ||
        std::fread(&damping, sizeof damping, 1, f) != 1) {
        std::fclose(f);
        return 2;
    }
    std::fclose(f);
    as_setup *h = as_setup_create_ex(
        vertices.data(), vertices.size() / 3, faces.data(), faces.size() / 3, face_material.data(), beta.data(),
        beta.size(), branches.empty() ? nullptr : branches.data(), branch_material.empty() ? nullptr : branch_material.data(),
        branch_material.size(), origin.data(), cell, u[0], u[1], u[2], source.data(), probes.data(), probes.size() / 3,
        1.0, u[3], damping, u[4], speed.empty() ? nullptr : speed.data(), flow_x.empty() ? nullptr : flow_x.data(),
        flow_y.empty() ? nullptr : flow_y.data(), u[5], 2, static_cast<unsigned>(u[6]));
    if (h == nullptr) {
        std::printf("refused: %s\n", as_setup_create_error());
        return 0;
    }
    uint64_t info[8] = {};
    as_setup_info(h, info);
    as_layout L;
    as_setup_layout(h, &L);
    std::printf("air %llu boundary %llu states %llu layer %llu probes %llu specials %llu links %llu taps %llu\n",
                static_cast<unsigned long long>(info[0]), static_cast<unsigned long long>(info[1]),
                static_cast<unsigned long long>(info[2]), static_cast<unsigned long long>(info[3]),
                static_cast<unsigned long long>(info[4]), static_cast<unsigned long long>(L.specials),
                static_cast<unsigned long long>(L.links), static_cast<unsigned long long>(L.taps));
    as_setup_destroy(h);
    return 0;
}
