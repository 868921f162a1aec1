// The host bottom-level build (neraulgraphics_amd/csrc/ngx_blas.cpp) driven over awkward meshes, for a build under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py): a single triangle, coincident and
// degenerate triangles, a fine grid, a random soup, several meshes at once with and without vertex normals. Each
// result is checked for what any correct build holds: every input triangle packed exactly once, children and
// triangle runs in range, a parent's box holding its children's, every skip escape in range. Exit status 0 when
// every check holds.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "ngx_host.h"

using namespace ngx::host;

namespace {

int failures = 0;

void check(const bool ok, const char *what, const int mesh_set) {
    if (!ok) {
        std::fprintf(stderr, "mesh set %d: %s\n", mesh_set, what);
        ++failures;
    }
}

struct Mesh {
    std::vector<float> v;
    std::vector<uint32_t> ix;
    std::vector<float> n;
};

Mesh grid(const int nx, const int ny, std::mt19937 &rng) {
    Mesh m;
    std::normal_distribution<float> z(0.0f, 0.05f);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) m.v.insert(m.v.end(), {(float)i / nx, (float)j / ny, z(rng)});
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const uint32_t a = j * (nx + 1) + i, b = a + 1, c = a + nx + 1, d = c + 1;
            m.ix.insert(m.ix.end(), {a, b, d, a, d, c});
        }
    return m;
}

Mesh soup(const int tris, std::mt19937 &rng, const bool degenerate) {
    Mesh m;
    std::uniform_real_distribution<float> u(-3.0f, 3.0f);
    for (int t = 0; t < tris; ++t) {
        const float x = u(rng), y = u(rng), z = u(rng);
        for (int k = 0; k < 3; ++k) {
            if (degenerate && t % 3 == 0) m.v.insert(m.v.end(), {x, y, z});  // every third triangle is a point
            else m.v.insert(m.v.end(), {x + 0.1f * u(rng), y + 0.1f * u(rng), z + 0.1f * u(rng)});
            m.ix.push_back(3 * t + k);
        }
    }
    return m;
}

Mesh coincident(const int copies) {
    Mesh m;
    m.v = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    for (int c = 0; c < copies; ++c) m.ix.insert(m.ix.end(), {0, 1, 2});
    return m;
}

void run(const std::vector<Mesh> &meshes, const int set, const bool normals) {
    std::vector<float> v, n;
    std::vector<uint32_t> ix;
    std::vector<int32_t> voff{0}, ioff{0};
    std::vector<uint8_t> has;
    for (const Mesh &m : meshes) {
        v.insert(v.end(), m.v.begin(), m.v.end());
        ix.insert(ix.end(), m.ix.begin(), m.ix.end());
        voff.push_back((int32_t)(v.size() / 3));
        ioff.push_back((int32_t)ix.size());
        for (size_t k = 0; k < m.v.size(); ++k) n.push_back(k % 3 == 2 ? 1.0f : 0.0f);
        has.push_back(normals && (has.size() % 2 == 0));
    }
    const MeshInput in{(int)meshes.size(), v.data(), voff.data(), ix.data(), ioff.data(), normals ? n.data() : nullptr,
                       normals ? has.data() : nullptr};
    BuiltMeshes out;
    build_meshes(in, out);
    const size_t total = ix.size() / 3;
    check(out.triangles.size() == 12 * total, "packed triangle count", set);
    check(out.tri_verts.size() == 3 * total, "triangle vertex count", set);
    // every input triangle exactly once: compare sorted vertex triples
    std::vector<std::array<int, 3>> want, got;
    for (size_t m = 0; m < meshes.size(); ++m)
        for (int32_t t = ioff[m] / 3; t < ioff[m + 1] / 3; ++t)
            want.push_back({voff[m] + (int)ix[3 * t], voff[m] + (int)ix[3 * t + 1], voff[m] + (int)ix[3 * t + 2]});
    for (size_t t = 0; t < total; ++t) got.push_back({out.tri_verts[3 * t], out.tri_verts[3 * t + 1], out.tri_verts[3 * t + 2]});
    std::sort(want.begin(), want.end());
    std::sort(got.begin(), got.end());
    check(want == got, "triangles packed once each", set);
    const int nodes = (int)out.nodes.size();
    for (int i = 0; i < nodes; ++i) {
        const ngx::BlasNode &b = out.nodes[i];
        if (b.count > 0) {
            check(b.count <= 4 && b.a >= 0 && (size_t)(b.a + b.count) <= total, "leaf run in range", set);
        } else {
            check(b.a > i && b.a + 1 < nodes, "children after their parent, in range", set);
            if (b.a > i && b.a + 1 < nodes)
                for (int c = 0; c < 2; ++c)
                    for (int k = 0; k < 3; ++k)
                        check(b.lo[k] <= out.nodes[b.a + c].lo[k] && b.hi[k] >= out.nodes[b.a + c].hi[k], "parent box holds child", set);
        }
        for (int o = 0; o < 8; ++o) check(out.skip[i].escape[o] >= -1 && out.skip[i].escape[o] < nodes, "skip escape in range", set);
    }
    check(out.mesh_roots.size() == meshes.size() && out.roots.size() == meshes.size(), "one root per mesh", set);
    check(normals == !out.attributes.empty(), "attributes only with normals", set);
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    run({coincident(1)}, 0, false);
    run({coincident(37)}, 1, false);
    run({soup(5, rng, true)}, 2, false);
    run({grid(40, 25, rng)}, 3, false);
    run({soup(3000, rng, true)}, 4, false);
    run({grid(3, 2, rng), coincident(9), soup(700, rng, false), grid(60, 40, rng)}, 5, true);
    run({soup(1, rng, true), soup(2, rng, false)}, 6, true);
    // a sorted-by-construction worst case for the partition: triangles along one line
    Mesh line;
    for (int t = 0; t < 4096; ++t) {
        line.v.insert(line.v.end(), {(float)t, 0, 0, (float)t + 0.5f, 1, 0, (float)t, 0, 1});
        line.ix.insert(line.ix.end(), {(uint32_t)(3 * t), (uint32_t)(3 * t + 1), (uint32_t)(3 * t + 2)});
    }
    run({line}, 7, false);
    std::printf("host bottom-level build: %d failed checks\n", failures);
    return failures == 0 ? 0 : 1;
}
