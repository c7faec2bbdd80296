// Synthetic unstructured hex meshes for the pressure-correction workloads.
//
// A structured lattice of candidate cells, optionally masked (backward-facing
// step), is emitted as a fully unstructured mesh in the reference's
// representation: face->vertex CSR, owner/neighbour per face (owner < neighbour
// for internal faces, -1 for boundary faces; reference mesh/ownership.h:25-129
// with the compact 32-bit widths of mesh/solver_mesh.h:11-15), cell->face CSR
// (mesh/cell.h:20-61) and a patch id per boundary face (mesh/boundary.h:31-40).
//
// Conventions (fixed here, stated in the design notes):
//  * internal faces are ordered by (owner, neighbour) in the final cell
//    numbering; boundary faces follow, grouped by patch, then by owner;
//  * face vertices wind owner -> neighbour (outward on the boundary);
//  * each cell lists its faces in ascending face id;
//  * cell ids are optionally shuffled with a seeded Fisher-Yates permutation;
//  * jitter moves interior lattice vertices by U(-a, a) times the smaller of
//    the two adjacent spacings along each axis; boundary vertices stay put.

#include "mesh_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace cfdxhip {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ^ 0x9E3779B97F4A7C15ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return next() % n; }
};

// Geometric wall clustering on [0, L] with n cells: `layers` cells at each end
// grow from h0 by ratio r, the core is uniform with spacing h0 * r^layers.
// r is found by bisection so the spacings sum to L.
std::vector<double> clustered_axis(uint32_t n, double L, uint32_t layers, double h0) {
    std::vector<double> x(n + 1, 0.0);
    if (layers == 0 || 2 * layers >= n) {
        for (uint32_t i = 0; i <= n; ++i) x[i] = L * double(i) / double(n);
        return x;
    }
    const uint32_t core = n - 2 * layers;
    auto total = [&](double r) {
        double s = 0.0, h = h0;
        for (uint32_t k = 0; k < layers; ++k) {
            s += h;
            h *= r;
        }
        return 2.0 * s + double(core) * h;
    };
    double lo = 1.0, hi = 4.0;
    if (total(lo) > L) throw std::invalid_argument("first cell too large for clustering");
    while (total(hi) < L) hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        (total(mid) < L ? lo : hi) = mid;
    }
    const double r = 0.5 * (lo + hi);
    std::vector<double> h(n);
    double hh = h0;
    for (uint32_t k = 0; k < layers; ++k) {
        h[k] = hh;
        h[n - 1 - k] = hh;
        hh *= r;
    }
    for (uint32_t k = layers; k < n - layers; ++k) h[k] = hh;
    for (uint32_t i = 0; i < n; ++i) x[i + 1] = x[i] + h[i];
    const double scale = L / x[n];
    for (auto& v : x) v *= scale;
    x[n] = L;
    return x;
}

// Geometric growth away from both ends with a fixed ratio, used for the
// wall-normal direction of the channel: spacing h0 * g^k until the spacing
// reaches the uniform core size, mirrored at the far end.
std::vector<double> graded_axis(uint32_t n, double L, double h0, double g) {
    if (g <= 1.0 || h0 <= 0.0) return clustered_axis(n, L, 0, 0.0);
    // Pick the number of graded layers so that the core spacing is at least
    // the last graded spacing, then solve the core size.
    uint32_t best_layers = 0;
    for (uint32_t layers = 1; 2 * layers < n; ++layers) {
        double s = 0.0, h = h0;
        for (uint32_t k = 0; k < layers; ++k) {
            s += h;
            h *= g;
        }
        const double core = (L - 2.0 * s) / double(n - 2 * layers);
        if (core <= 0.0) break;
        if (core >= h / g) best_layers = layers;
        else break;
    }
    std::vector<double> hv(n);
    double s = 0.0, h = h0;
    for (uint32_t k = 0; k < best_layers; ++k) {
        hv[k] = hv[n - 1 - k] = h;
        s += h;
        h *= g;
    }
    const double core = (L - 2.0 * s) / double(n - 2 * best_layers);
    for (uint32_t k = best_layers; k < n - best_layers; ++k) hv[k] = core;
    std::vector<double> x(n + 1, 0.0);
    for (uint32_t i = 0; i < n; ++i) x[i + 1] = x[i] + hv[i];
    x[n] = L;
    return x;
}

}  // namespace

HostMesh generate_mesh(const MeshSpec& s) {
    if (s.nx == 0 || s.ny == 0 || s.nz == 0) throw std::invalid_argument("empty lattice");
    const uint32_t nx = s.nx, ny = s.ny, nz = s.nz;
    HostMesh m;
    // ---- axes
    std::vector<double> X, Y, Z;
    if (s.kind == MESH_CHANNEL_STEP) {
        X = clustered_axis(nx, s.lx, 0, 0.0);
        const double h_stream = s.lx / double(nx);
        const double h0 = s.first_aspect > 0.0 ? h_stream / s.first_aspect : 0.0;
        Y = graded_axis(ny, s.ly, h0, s.growth);
        Z = clustered_axis(nz, s.lz, 0, 0.0);
    } else {
        X = clustered_axis(nx, s.lx, 0, 0.0);
        const double h0 = s.first_aspect > 0.0 ? (s.lx / double(nx)) / s.first_aspect : 0.0;
        Y = clustered_axis(ny, s.ly, s.wall_layers, h0);
        Z = clustered_axis(nz, s.lz, 0, 0.0);
    }
    // ---- active-cell mask (step: lower-left block removed)
    const uint64_t ncand = uint64_t(nx) * ny * nz;
    auto sid = [&](uint32_t i, uint32_t j, uint32_t k) { return (uint64_t(k) * ny + j) * nx + i; };
    std::vector<uint8_t> active(ncand, 1);
    uint32_t step_i = 0, step_j = 0;
    if (s.kind == MESH_CHANNEL_STEP) {
        step_i = uint32_t(std::lround(s.step_length_fraction * nx));
        const double yh = s.step_height_fraction * s.ly;
        step_j = 0;
        while (step_j < ny && Y[step_j] < yh - 1e-12 * s.ly) ++step_j;
        for (uint32_t k = 0; k < nz; ++k)
            for (uint32_t j = 0; j < step_j; ++j)
                for (uint32_t i = 0; i < step_i; ++i) active[sid(i, j, k)] = 0;
    }
    auto is_active = [&](int64_t i, int64_t j, int64_t k) {
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return false;
        return active[sid(uint32_t(i), uint32_t(j), uint32_t(k))] != 0;
    };
    // ---- cell numbering (structured order, then optional shuffle)
    std::vector<uint32_t> cell_of(ncand, UINT32_MAX);
    uint32_t nc = 0;
    for (uint64_t c = 0; c < ncand; ++c)
        if (active[c]) cell_of[c] = nc++;
    std::vector<uint32_t> perm(nc);
    std::iota(perm.begin(), perm.end(), 0u);
    if (s.permute) {
        Rng rng(s.seed * 2654435761ull + 17);
        for (uint32_t i = nc; i > 1; --i) {
            const uint32_t j = uint32_t(rng.below(i));
            std::swap(perm[i - 1], perm[j]);
        }
    }
    for (uint64_t c = 0; c < ncand; ++c)
        if (cell_of[c] != UINT32_MAX) cell_of[c] = perm[cell_of[c]];
    // ---- points (full lattice), jitter on interior vertices
    const uint64_t npx = nx + 1, npy = ny + 1, npz = nz + 1;
    const uint64_t np = npx * npy * npz;
    auto pid = [&](uint32_t i, uint32_t j, uint32_t k) {
        return uint32_t((uint64_t(k) * npy + j) * npx + i);
    };
    m.px.resize(np);
    m.py.resize(np);
    m.pz.resize(np);
    {
        Rng rng(s.seed * 0x2545F4914F6CDD1Dull + 99);
        for (uint32_t k = 0; k <= nz; ++k)
            for (uint32_t j = 0; j <= ny; ++j)
                for (uint32_t i = 0; i <= nx; ++i) {
                    const uint32_t p = pid(i, j, k);
                    m.px[p] = X[i];
                    m.py[p] = Y[j];
                    m.pz[p] = Z[k];
                    if (s.jitter <= 0.0) continue;
                    // A vertex is interior when all eight lattice cells around it are active.
                    bool interior = true;
                    for (int dk = -1; dk <= 0 && interior; ++dk)
                        for (int dj = -1; dj <= 0 && interior; ++dj)
                            for (int di = -1; di <= 0 && interior; ++di)
                                interior = is_active(int64_t(i) + di, int64_t(j) + dj, int64_t(k) + dk);
                    const double ux = rng.uniform(), uy = rng.uniform(), uz = rng.uniform();
                    if (!interior) continue;
                    const double hx = std::min(X[i] - X[i - 1], X[i + 1] - X[i]);
                    const double hy = std::min(Y[j] - Y[j - 1], Y[j + 1] - Y[j]);
                    const double hz = std::min(Z[k] - Z[k - 1], Z[k + 1] - Z[k]);
                    m.px[p] += (2.0 * ux - 1.0) * s.jitter * hx;
                    m.py[p] += (2.0 * uy - 1.0) * s.jitter * hy;
                    if (s.kind != MESH_EXTRUDED_2D) m.pz[p] += (2.0 * uz - 1.0) * s.jitter * hz;
                }
    }
    // ---- faces
    struct RawFace {
        uint32_t owner;
        int32_t neigh;
        uint32_t patch;
        uint32_t v[4];
    };
    std::vector<RawFace> internal, boundary;
    internal.reserve(3 * uint64_t(nc));
    // Patch ids: 0 x-min, 1 x-max, 2 y-min, 3 y-max, 4 z-min, 5 z-max, 6 step walls.
    auto emit = [&](int axis, uint32_t i, uint32_t j, uint32_t k) {
        // Face on lattice plane `axis` at index (i,j,k): separates lattice
        // cell (i,j,k) - e_axis (minus side) from (i,j,k) (plus side).
        uint32_t v[4];
        if (axis == 0) {
            v[0] = pid(i, j, k); v[1] = pid(i, j + 1, k); v[2] = pid(i, j + 1, k + 1); v[3] = pid(i, j, k + 1);
        } else if (axis == 1) {
            v[0] = pid(i, j, k); v[1] = pid(i, j, k + 1); v[2] = pid(i + 1, j, k + 1); v[3] = pid(i + 1, j, k);
        } else {
            v[0] = pid(i, j, k); v[1] = pid(i + 1, j, k); v[2] = pid(i + 1, j + 1, k); v[3] = pid(i, j + 1, k);
        }
        // Winding above gives a normal along +axis (minus -> plus side).
        const int64_t mi = int64_t(i) - (axis == 0), mj = int64_t(j) - (axis == 1), mk = int64_t(k) - (axis == 2);
        const bool am = is_active(mi, mj, mk), ap = is_active(i, j, k);
        if (!am && !ap) return;
        auto reversed = [&](RawFace& f) { std::swap(f.v[1], f.v[3]); };
        RawFace f;
        std::memcpy(f.v, v, sizeof v);
        if (am && ap) {
            const uint32_t cm = cell_of[sid(uint32_t(mi), uint32_t(mj), uint32_t(mk))];
            const uint32_t cp = cell_of[sid(i, j, k)];
            f.owner = std::min(cm, cp);
            f.neigh = int32_t(std::max(cm, cp));
            f.patch = UINT32_MAX;
            if (f.owner == cp) reversed(f);  // wind owner -> neighbour
            internal.push_back(f);
            return;
        }
        const bool domain_edge = (axis == 0 && (i == 0 || i == nx)) || (axis == 1 && (j == 0 || j == ny)) ||
                                 (axis == 2 && (k == 0 || k == nz));
        if (ap) {  // active cell on the plus side: outward normal is -axis
            f.owner = cell_of[sid(i, j, k)];
            reversed(f);
            f.patch = domain_edge ? uint32_t(2 * axis) : 6u;
        } else {
            f.owner = cell_of[sid(uint32_t(mi), uint32_t(mj), uint32_t(mk))];
            f.patch = domain_edge ? uint32_t(2 * axis + 1) : 6u;
        }
        f.neigh = -1;
        boundary.push_back(f);
    };
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i <= nx; ++i) emit(0, i, j, k);
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j <= ny; ++j)
            for (uint32_t i = 0; i < nx; ++i) emit(1, i, j, k);
    for (uint32_t k = 0; k <= nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i) emit(2, i, j, k);
    // Internal faces by (owner, neighbour): counting sort on owner, then
    // insertion sort of the (at most six) faces of each owner by neighbour.
    {
        std::vector<uint32_t> cnt(uint64_t(nc) + 1, 0);
        for (const auto& f : internal) ++cnt[f.owner + 1];
        for (uint32_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
        std::vector<RawFace> sorted(internal.size());
        std::vector<uint32_t> cur(cnt.begin(), cnt.end() - 1);
        for (const auto& f : internal) sorted[cur[f.owner]++] = f;
        for (uint32_t c = 0; c < nc; ++c)
            std::sort(sorted.begin() + cnt[c], sorted.begin() + cnt[c + 1],
                      [](const RawFace& a, const RawFace& b) { return a.neigh < b.neigh; });
        internal.swap(sorted);
    }
    std::stable_sort(boundary.begin(), boundary.end(), [](const RawFace& a, const RawFace& b) {
        if (a.patch != b.patch) return a.patch < b.patch;
        return a.owner < b.owner;
    });
    const uint64_t nf = internal.size() + boundary.size();
    if (nf >= UINT32_MAX || 4 * nf >= UINT32_MAX) throw std::invalid_argument("mesh too large for 32-bit faces");
    m.n_cells = nc;
    m.n_internal_faces = internal.size();
    m.foff.resize(nf + 1);
    m.fvtx.resize(4 * nf);
    m.owner.resize(nf);
    m.neigh.resize(nf);
    m.patch_of_face.resize(nf);
    uint64_t f = 0;
    for (const auto* list : {&internal, &boundary})
        for (const auto& rf : *list) {
            m.foff[f] = uint32_t(4 * f);
            std::memcpy(&m.fvtx[4 * f], rf.v, sizeof rf.v);
            m.owner[f] = rf.owner;
            m.neigh[f] = rf.neigh;
            m.patch_of_face[f] = rf.patch;
            ++f;
        }
    m.foff[nf] = uint32_t(4 * nf);
    internal.clear();
    internal.shrink_to_fit();
    boundary.clear();
    boundary.shrink_to_fit();
    // ---- cell -> faces, ascending face id (faces are visited in id order)
    m.coff.assign(uint64_t(nc) + 1, 0);
    for (uint64_t g = 0; g < nf; ++g) {
        ++m.coff[m.owner[g] + 1];
        if (m.neigh[g] >= 0) ++m.coff[uint32_t(m.neigh[g]) + 1];
    }
    for (uint32_t c = 0; c < nc; ++c) m.coff[c + 1] += m.coff[c];
    m.cfaces.resize(m.coff[nc]);
    {
        std::vector<uint32_t> cur(m.coff.begin(), m.coff.end() - 1);
        for (uint64_t g = 0; g < nf; ++g) {
            m.cfaces[cur[m.owner[g]]++] = uint32_t(g);
            if (m.neigh[g] >= 0) m.cfaces[cur[uint32_t(m.neigh[g])]++] = uint32_t(g);
        }
    }
    // ---- patches
    m.n_patches = 7;
    m.patch_kind.assign(7, PATCH_WALL);
    if (s.kind == MESH_EXTRUDED_2D) m.patch_kind[4] = m.patch_kind[5] = PATCH_EMPTY;
    // Structured coordinates of every cell (for seeded fields / smoothing).
    m.cell_ijk.resize(3 * uint64_t(nc));
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i) {
                const uint32_t c = cell_of[sid(i, j, k)];
                if (c == UINT32_MAX) continue;
                m.cell_ijk[3 * uint64_t(c)] = i;
                m.cell_ijk[3 * uint64_t(c) + 1] = j;
                m.cell_ijk[3 * uint64_t(c) + 2] = k;
            }
    return m;
}

}  // namespace cfdxhip
