// Element lists in su2 form: C ABI, su2 text reader/writer, synthetic mesh generators.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <unordered_map>

#include "gr_internal.hpp"

namespace gr {

bool elem_kind(uint32_t kind, ElemKind &out) {
    // Same local face tables as the reference (src/meshio.cu:315-475); they decide the slot
    // order of every cell and so must be identical.
    static const ElemKind TRI = {3, 1, {{0, 1, 2, -1}}};
    static const ElemKind QUAD = {4, 1, {{0, 1, 2, 3}}};
    static const ElemKind TET = {4, 4, {{0, 2, 1, -1}, {0, 1, 3, -1}, {1, 2, 3, -1}, {0, 3, 2, -1}}};
    static const ElemKind HEX = {8, 6, {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7},
                                        {3, 2, 1, 0}, {4, 5, 6, 7}}};
    static const ElemKind WEDGE = {6, 5, {{0, 1, 2, -1}, {3, 5, 4, -1}, {0, 2, 5, 3}, {1, 4, 5, 2},
                                          {1, 0, 3, 4}}};
    static const ElemKind PYR = {5, 5, {{0, 1, 2, 3}, {0, 4, 3, -1}, {0, 1, 4, -1}, {1, 2, 4, -1},
                                        {2, 3, 4, -1}}};
    switch (kind) {
    case GR_TRI: out = TRI; return true;
    case GR_QUAD: out = QUAD; return true;
    case GR_TET: out = TET; return true;
    case GR_HEX: out = HEX; return true;
    case GR_WEDGE: out = WEDGE; return true;
    case GR_PYRAMID: out = PYR; return true;
    default: return false;
    }
}

static void check_elements(uint32_t n_nodes, uint32_t n, const uint32_t *kinds,
                           const uint32_t *starts, const uint32_t *nodes, bool surface) {
    require(starts[0] == 0, "element starts must begin at 0");
    for (uint32_t i = 0; i < n; ++i) {
        ElemKind k;
        require(elem_kind(kinds[i], k), "unknown element kind " + std::to_string(kinds[i]));
        bool is_surface = kinds[i] == GR_TRI || kinds[i] == GR_QUAD;
        require(is_surface == surface, surface ? "marker elements must be triangles or quads"
                                               : "volume elements must be tet/hex/wedge/pyramid");
        require(starts[i + 1] >= starts[i] && starts[i + 1] - starts[i] == (uint32_t)k.n_nodes,
                "element " + std::to_string(i) + " has the wrong node count for its kind");
        for (uint32_t q = starts[i]; q < starts[i + 1]; ++q)
            require(nodes[q] < n_nodes, "element node index out of range");
    }
}

// ---- boundary extraction used by the generators ----------------------------------------

struct FaceKey {
    uint32_t k[4];
    bool operator==(const FaceKey &o) const { return std::memcmp(k, o.k, sizeof k) == 0; }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey &f) const {
        uint64_t h = 1469598103934665603ull;
        for (uint32_t v : f.k) { h ^= v; h *= 1099511628211ull; }
        return (size_t)h;
    }
};

static FaceKey sorted_key(const uint32_t *n, int cnt) {
    FaceKey key;
    for (int i = 0; i < 4; ++i) key.k[i] = i < cnt ? n[i] : 0xFFFFFFFFu;
    std::sort(key.k, key.k + 4);
    return key;
}

// Adds one marker per name; classify(face nodes, count) returns the marker index of a
// boundary face (faces that belong to a single element).  Faces are emitted in element order.
template <typename Classify>
static void add_boundary_markers(gr_elements &e, const std::vector<std::string> &names,
                                 Classify classify) {
    std::unordered_map<FaceKey, uint32_t, FaceKeyHash> count;
    count.reserve(e.kinds.size() * 6);
    for (uint32_t el = 0; el < e.n_elems(); ++el) {
        ElemKind k;
        elem_kind(e.kinds[el], k);
        for (int f = 0; f < k.n_faces; ++f) {
            uint32_t fn[4];
            int c = 0;
            for (int j = 0; j < 4; ++j)
                if (k.f[f][j] >= 0) fn[c++] = e.nodes[e.starts[el] + k.f[f][j]];
            count[sorted_key(fn, c)]++;
        }
    }
    e.markers.clear();
    for (const auto &n : names) {
        Marker m;
        m.name = n;
        m.starts.push_back(0);
        e.markers.push_back(m);
    }
    for (uint32_t el = 0; el < e.n_elems(); ++el) {
        ElemKind k;
        elem_kind(e.kinds[el], k);
        for (int f = 0; f < k.n_faces; ++f) {
            uint32_t fn[4];
            int c = 0;
            for (int j = 0; j < 4; ++j)
                if (k.f[f][j] >= 0) fn[c++] = e.nodes[e.starts[el] + k.f[f][j]];
            if (count[sorted_key(fn, c)] != 1) continue;
            int mk = classify(fn, c);
            require(mk >= 0 && mk < (int)names.size(), "generator: unclassified boundary face");
            Marker &m = e.markers[mk];
            m.kinds.push_back(c == 3 ? GR_TRI : GR_QUAD);
            for (int j = 0; j < c; ++j) m.nodes.push_back(fn[j]);
            m.starts.push_back((uint32_t)m.nodes.size());
        }
    }
}

// Uniform double in [0,1) from the top 53 bits of a 64-bit draw (platform independent).
static double u01(std::mt19937_64 &g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }

void gen_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, int tets,
             double jitter, uint64_t seed, gr_elements &e) {
    require(nx > 0 && ny > 0 && nz > 0, "box: cell counts must be positive");
    require(lx > 0 && ly > 0 && lz > 0, "box: lengths must be positive");
    require(jitter >= 0 && jitter < 0.25, "box: jitter must be in [0, 0.25)");
    const uint64_t nn = (uint64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const uint64_t ne = (uint64_t)nx * ny * nz * (tets ? 6 : 1);
    require(nn < 0xFFFFFFF0ull && ne * 8 < 0xFFFFFFF0ull, "box: mesh too large for 32-bit indices");
    auto nid = [&](uint32_t i, uint32_t j, uint32_t k) { return i + (nx + 1) * (j + (ny + 1) * k); };
    const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
    std::mt19937_64 rng(seed);
    e = gr_elements();
    e.xyz.resize(3 * nn);
    for (uint32_t k = 0; k <= nz; ++k)
        for (uint32_t j = 0; j <= ny; ++j)
            for (uint32_t i = 0; i <= nx; ++i) {
                double x = i * hx, y = j * hy, z = k * hz;
                bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                if (interior && jitter > 0) {
                    x += (2.0 * u01(rng) - 1.0) * jitter * hx;
                    y += (2.0 * u01(rng) - 1.0) * jitter * hy;
                    z += (2.0 * u01(rng) - 1.0) * jitter * hz;
                }
                uint32_t n = nid(i, j, k);
                e.xyz[3 * n] = x;
                e.xyz[3 * n + 1] = y;
                e.xyz[3 * n + 2] = z;
            }
    e.starts.push_back(0);
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i) {
                // hex corners in su2 order
                uint32_t h[8] = {nid(i, j, k),         nid(i + 1, j, k),     nid(i + 1, j + 1, k),
                                 nid(i, j + 1, k),     nid(i, j, k + 1),     nid(i + 1, j, k + 1),
                                 nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)};
                if (!tets) {
                    e.kinds.push_back(GR_HEX);
                    e.nodes.insert(e.nodes.end(), h, h + 8);
                    e.starts.push_back((uint32_t)e.nodes.size());
                    continue;
                }
                // Kuhn split along the diagonal h[0]-h[6]: one tet per axis permutation,
                // path 0 -> e_a -> e_a+e_b -> (1,1,1).  Corner of offset (a,b,c) is
                // corner[a + 2b + 4c] below.
                uint32_t corner[8] = {h[0], h[1], h[3], h[2], h[4], h[5], h[7], h[6]};
                static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                                {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                for (const auto &p : perms) {
                    int o1 = 1 << p[0], o2 = o1 | (1 << p[1]);
                    uint32_t t[4] = {corner[0], corner[o1], corner[o2], corner[7]};
                    // orient so that the su2 volume is positive
                    const double *a = &e.xyz[3 * t[0]], *b = &e.xyz[3 * t[1]];
                    const double *c = &e.xyz[3 * t[2]], *d = &e.xyz[3 * t[3]];
                    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                    double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                    double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
                    double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                                 u[2] * (v[0] * w[1] - v[1] * w[0]);
                    if (det < 0) std::swap(t[1], t[2]);
                    e.kinds.push_back(GR_TET);
                    e.nodes.insert(e.nodes.end(), t, t + 4);
                    e.starts.push_back((uint32_t)e.nodes.size());
                }
            }
    const uint32_t sx = nx + 1, sxy = (nx + 1) * (ny + 1);
    add_boundary_markers(e, {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"},
                         [&](const uint32_t *fn, int c) {
                             auto all = [&](auto pred) {
                                 for (int q = 0; q < c; ++q)
                                     if (!pred(fn[q])) return false;
                                 return true;
                             };
                             if (all([&](uint32_t n) { return n % sx == 0; })) return 0;
                             if (all([&](uint32_t n) { return n % sx == nx; })) return 1;
                             if (all([&](uint32_t n) { return (n % sxy) / sx == 0; })) return 2;
                             if (all([&](uint32_t n) { return (n % sxy) / sx == ny; })) return 3;
                             if (all([&](uint32_t n) { return n / sxy == 0; })) return 4;
                             if (all([&](uint32_t n) { return n / sxy == nz; })) return 5;
                             return -1;
                         });
}

void gen_ogrid(uint32_t nt, uint32_t nr, uint32_t nz, double r0, double r1, double ratio,
               double dz, gr_elements &e) {
    require(nt >= 3 && nr > 0 && nz > 0, "ogrid: need n_theta >= 3, n_r > 0, n_z > 0");
    require(r0 > 0 && r1 > r0 && ratio > 0 && dz > 0, "ogrid: bad radii, ratio or dz");
    const uint64_t nn = (uint64_t)nt * (nr + 1) * (nz + 1);
    require(nn < 0xFFFFFFF0ull && (uint64_t)nt * nr * nz * 8 < 0xFFFFFFF0ull, "ogrid: too large");
    std::vector<double> r(nr + 1);
    if (std::fabs(ratio - 1.0) < 1e-14) {
        for (uint32_t j = 0; j <= nr; ++j) r[j] = r0 + (r1 - r0) * j / nr;
    } else {
        double d0 = (r1 - r0) * (ratio - 1.0) / (std::pow(ratio, (double)nr) - 1.0);
        for (uint32_t j = 0; j <= nr; ++j) r[j] = r0 + d0 * (std::pow(ratio, (double)j) - 1.0) / (ratio - 1.0);
        r[nr] = r1;
    }
    auto nid = [&](uint32_t t, uint32_t j, uint32_t k) { return (t % nt) + nt * (j + (nr + 1) * k); };
    e = gr_elements();
    e.xyz.resize(3 * nn);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t k = 0; k <= nz; ++k)
        for (uint32_t j = 0; j <= nr; ++j)
            for (uint32_t t = 0; t < nt; ++t) {
                double th = two_pi * t / nt;
                uint32_t n = nid(t, j, k);
                e.xyz[3 * n] = r[j] * std::cos(th);
                e.xyz[3 * n + 1] = r[j] * std::sin(th);
                e.xyz[3 * n + 2] = k * dz;
            }
    e.starts.push_back(0);
    for (uint32_t k = 0; k < nz; ++k)
        for (uint32_t j = 0; j < nr; ++j)
            for (uint32_t t = 0; t < nt; ++t) {
                uint32_t h[8] = {nid(t, j, k),         nid(t + 1, j, k),     nid(t + 1, j + 1, k),
                                 nid(t, j + 1, k),     nid(t, j, k + 1),     nid(t + 1, j, k + 1),
                                 nid(t + 1, j + 1, k + 1), nid(t, j + 1, k + 1)};
                e.kinds.push_back(GR_HEX);
                e.nodes.insert(e.nodes.end(), h, h + 8);
                e.starts.push_back((uint32_t)e.nodes.size());
            }
    const uint32_t per_layer = nt * (nr + 1);
    add_boundary_markers(e, {"wall", "farfield", "zmin", "zmax"}, [&](const uint32_t *fn, int c) {
        auto all = [&](auto pred) {
            for (int q = 0; q < c; ++q)
                if (!pred(fn[q])) return false;
            return true;
        };
        if (all([&](uint32_t n) { return n / per_layer == 0; })) return 2;
        if (all([&](uint32_t n) { return n / per_layer == nz; })) return 3;
        if (all([&](uint32_t n) { return (n % per_layer) / nt == 0; })) return 0;
        if (all([&](uint32_t n) { return (n % per_layer) / nt == nr; })) return 1;
        return -1;
    });
}

// NACA 4-digit symmetric thickness, closed trailing edge.
static double naca_half_thickness(double x, double t) {
    return 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
}

// Ratio r with h0 * (r^n - 1)/(r - 1) = 1 (geometric spacing from h0 over unit length).
static double stretch_ratio(double h0, uint32_t n) {
    if (h0 * n >= 1.0) return 1.0;
    double lo = 1.0 + 1e-12, hi = 4.0;
    for (int it = 0; it < 200; ++it) {
        double r = 0.5 * (lo + hi);
        double len = h0 * (std::pow(r, (double)n) - 1.0) / (r - 1.0);
        (len > 1.0 ? hi : lo) = r;
    }
    return 0.5 * (lo + hi);
}

void gen_cgrid(uint32_t ni, uint32_t nj, uint32_t nwake, double thick, double radius, double h0, double dz,
               gr_elements &e) {
    require(ni >= 16 && nj >= 2 && nwake >= 2 && 2 * nwake + 8 <= ni && (ni - 2 * nwake) % 2 == 0,
            "cgrid: need ni >= 2*n_wake + 8 with an even airfoil count");
    require(thick > 0 && thick < 0.5 && radius > 2 && h0 > 0 && h0 < 0.5 && dz > 0, "cgrid: bad parameters");
    const uint32_t naf = ni - 2 * nwake, nhalf = naf / 2;
    const double pi = 3.14159265358979323846;
    // inner line (j = 0): lower wake (outflow -> TE), lower surface (TE -> LE), upper surface
    // (LE -> TE), upper wake (TE -> outflow).  Outer line (j = nj): the matching C boundary.
    std::vector<double> px(ni + 1), py(ni + 1), qx(ni + 1), qy(ni + 1);
    const double x_out = 1.0 + radius;
    const double te_spacing = 0.5 * (1 - std::cos(pi / nhalf)) ;   // ~ chordwise spacing at the TE
    const double wr = stretch_ratio(std::max(te_spacing, 1e-4) / radius, nwake);
    std::vector<double> wx(nwake + 1);
    {
        double acc = 0, h = std::max(te_spacing, 1e-4) / radius;
        wx[0] = 0;
        for (uint32_t k = 1; k <= nwake; ++k) {
            acc += h;
            h *= wr;
            wx[k] = acc;
        }
        for (uint32_t k = 0; k <= nwake; ++k) wx[k] = 1.0 + radius * wx[k] / wx[nwake];
    }
    for (uint32_t i = 0; i <= ni; ++i) {
        if (i <= nwake) {                       // lower wake, x decreasing to the TE
            px[i] = wx[nwake - i];
            py[i] = 0;
            qx[i] = px[i];
            qy[i] = -radius;
        } else if (i >= ni - nwake) {           // upper wake
            px[i] = wx[i - (ni - nwake)];
            py[i] = 0;
            qx[i] = px[i];
            qy[i] = radius;
        } else {                                // airfoil, cosine spacing; beta: 0 at TE lower .. pi at LE .. 2pi at TE upper
            const double beta = pi * (double)(i - nwake) / nhalf;
            const double x = 0.5 * (1 + std::cos(beta));
            const double yt = naca_half_thickness(std::min(1.0, std::max(0.0, x)), thick);
            px[i] = x;
            py[i] = (i - nwake < nhalf) ? -yt : yt;
            // outer C: half circle of the same radius around (1, 0), angle matching beta
            const double phi = -0.5 * pi - pi * (double)(i - nwake) / naf;   // -pi/2 .. -3pi/2
            qx[i] = 1.0 + radius * std::cos(phi);
            qy[i] = radius * std::sin(phi);
        }
    }
    px[nwake] = 1.0; py[nwake] = 0.0;           // trailing edge shared by surface and wake
    px[ni - nwake] = 1.0; py[ni - nwake] = 0.0;
    const double rj = stretch_ratio(h0 / radius, nj);
    std::vector<double> s(nj + 1);
    {
        double acc = 0, h = h0 / radius;
        s[0] = 0;
        for (uint32_t j = 1; j <= nj; ++j) {
            acc += h;
            h *= rj;
            s[j] = acc;
        }
        for (uint32_t j = 0; j <= nj; ++j) s[j] /= s[nj];
    }
    // node ids; the upper wake's inner nodes are the lower wake's (wake cut)
    const uint32_t per_layer = (ni + 1) * (nj + 1);
    auto raw = [&](uint32_t i, uint32_t j, uint32_t k) { return i + (ni + 1) * (j + (nj + 1) * k); };
    std::vector<uint32_t> remap(2 * per_layer);
    std::vector<uint32_t> node_i(2 * per_layer), node_j(2 * per_layer), node_k(2 * per_layer);
    uint32_t next = 0;
    std::vector<double> xyz;
    for (uint32_t k = 0; k < 2; ++k)
        for (uint32_t j = 0; j <= nj; ++j)
            for (uint32_t i = 0; i <= ni; ++i) {
                uint32_t r = raw(i, j, k);
                if (j == 0 && i >= ni - nwake) {
                    remap[r] = remap[raw(ni - i, 0, k)];
                    continue;
                }
                remap[r] = next;
                node_i[next] = i;
                node_j[next] = j;
                node_k[next] = k;
                ++next;
                xyz.push_back(px[i] + s[j] * (qx[i] - px[i]));
                xyz.push_back(py[i] + s[j] * (qy[i] - py[i]));
                xyz.push_back(k * dz);
            }
    e = gr_elements();
    e.xyz = std::move(xyz);
    e.starts.push_back(0);
    for (uint32_t j = 0; j < nj; ++j)
        for (uint32_t i = 0; i < ni; ++i) {
            uint32_t h[8] = {remap[raw(i, j, 0)],     remap[raw(i + 1, j, 0)], remap[raw(i + 1, j + 1, 0)],
                             remap[raw(i, j + 1, 0)], remap[raw(i, j, 1)],     remap[raw(i + 1, j, 1)],
                             remap[raw(i + 1, j + 1, 1)], remap[raw(i, j + 1, 1)]};
            e.kinds.push_back(GR_HEX);
            e.nodes.insert(e.nodes.end(), h, h + 8);
            e.starts.push_back((uint32_t)e.nodes.size());
        }
    add_boundary_markers(e, {"airfoil", "farfield", "zmin", "zmax"}, [&](const uint32_t *fn, int c) {
        auto all = [&](auto pred) {
            for (int q = 0; q < c; ++q)
                if (!pred(fn[q])) return false;
            return true;
        };
        if (all([&](uint32_t n) { return node_k[n] == 0; })) return 2;
        if (all([&](uint32_t n) { return node_k[n] == 1; })) return 3;
        if (all([&](uint32_t n) { return node_j[n] == 0; })) return 0;
        return 1;   // outer C boundary and the two outflow planes
    });
}

// ---- swept wing in a half-sphere: prism layers + tetrahedra ----------------------------------
//
// Structured O-grid topology (i around the section, j outward, k along the span), z spanwise
// from the symmetry plane z = 0.  Stations k <= n_span carry the tapered, swept NACA section;
// above the tip the section closes to a zero-thickness plate whose upper and lower nodes are
// merged, capping the tip with a low ridge.  Cells: hexes j < n_prism split into two prisms
// along the constant-j face diagonal (i,k)-(i+1,k+1); hexes j >= n_prism split into six
// tetrahedra (Kuhn, main diagonal (i,j,k)-(i+1,j+1,k+1), which cuts every constant-j face
// along the same diagonal).

namespace {

struct V3 {
    double x, y, z;
};

V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

double det3(const V3 &u, const V3 &v, const V3 &w) {
    return u.x * (v.y * w.z - v.z * w.y) - u.y * (v.x * w.z - v.z * w.x) + u.z * (v.x * w.y - v.y * w.x);
}

}  // namespace

void gen_wing(const gr_wing_params &p, gr_elements &e) {
    const uint32_t ni = p.n_around, np = p.n_prism, nt = p.n_tet, ns = p.n_span, nc = p.n_cap;
    require(ni >= 8 && ni % 2 == 0, "wing: n_around must be even and >= 8");
    require(np >= 1 && nt >= 1 && ns >= 2 && nc >= 2, "wing: need n_prism, n_tet >= 1 and n_span, n_cap >= 2");
    require(p.thickness > 0 && p.thickness < 0.5 && p.taper > 0 && p.taper <= 1 && p.semispan > 0,
            "wing: bad thickness, taper or semispan");
    require(p.sweep_deg > -60 && p.sweep_deg < 60, "wing: sweep must be within +-60 degrees");
    require(p.first_height > 0 && p.growth >= 1 && p.radius > 4 * (p.semispan + 1), "wing: bad layers or radius");
    require(p.scale > 0 && std::isfinite(p.scale), "wing: scale must be positive");
    const uint32_t nj = np + nt, K = ns + nc;
    // Stations above the tip over which the section's thickness tapers to the zero-thickness
    // plate (set with the cap heights below): the tip is capped by a short wedge, about twice as
    // tall as the tip section is thick, rather than one steep ridge, whose faces (the section at
    // the tip station against the plate one station up) twist strongly where the thickness
    // changes fast, at the leading edge.
    uint32_t n_taper = 1;
    const uint64_t n_grid = (uint64_t)ni * (nj + 1) * (K + 1);
    require(n_grid * 8 < 0xFFFFFFF0ull, "wing: mesh too large for 32-bit indices");
    const double pi = 3.14159265358979323846;
    const double b = p.semispan, R = p.radius, tsw = std::tan(p.sweep_deg * pi / 180.0);
    const double cx0 = 0.5, cy0 = 0.0;   // axis of the half-sphere (root mid-chord)
    // prism layer offsets
    std::vector<double> d(np + 1, 0.0);
    {
        double h = p.first_height;
        for (uint32_t j = 1; j <= np; ++j) {
            d[j] = d[j - 1] + h;
            h *= p.growth;
        }
    }
    const double h_tet = p.first_height * std::pow(p.growth, (double)np);
    // unit section, ring order TE -> lower -> LE -> upper
    std::vector<double> ux(ni), uy(ni), th(ni), beta_of(ni);
    for (uint32_t i = 0; i < ni; ++i) {
        // cosine spacing relaxed at the leading and trailing edges (beta + sin(2 beta)/4)
        const double beta0 = 2 * pi * i / ni, beta = beta0 + 0.25 * std::sin(2 * beta0);
        const double X = 0.5 * (1 + std::cos(beta));
        const double t = naca_half_thickness(std::min(1.0, std::max(0.0, X)), p.thickness);
        ux[i] = X;
        uy[i] = (i == 0 || 2 * i == ni) ? 0.0 : (2 * i < ni ? -t : t);
        th[i] = -beta0;
        beta_of[i] = beta;
    }
    auto nid = [&](uint32_t i, uint32_t j, uint32_t k) { return (i % ni) + ni * (j + (nj + 1) * k); };
    std::vector<V3> X(n_grid);
    // wing stations
    // Far-field latitudes: the wing stations' radial lines end on the sphere at latitudes
    // spread evenly over [0, phi_b] (not at the station's own height), so cells near the far
    // field are not flattened to the station spacing; the stations above the tip continue at
    // the same latitude step up to phi_top.  Station heights are half uniform, half clustered
    // towards the tip, so the tip spacing stays within a factor two of the mean.  Above phi_top
    // the dome holds the top layout of every ring (below); at 70 degrees its disc has a radius
    // of 0.34 R, room for rings a fair fraction of the cap's top steps apart (at 85 degrees, 0.09
    // R, the column's top cells were slivers with face-area ratios above 1e5, whose residual
    // oscillated without end and held the solve at a drop of 1e-2).
    const double phi_b = 60.0 * pi / 180.0;
    const double phi_top = std::min(70.0 * pi / 180.0, phi_b * (1.0 + (double)nc / ns));
    require(R * std::sin(phi_b) > 2 * b, "wing: far field too close to the tip");
    auto station_z = [&](uint32_t k) { return b * (0.5 * k / ns + 0.5 * std::sin(0.5 * pi * k / ns)); };
    for (uint32_t k = 0; k <= ns; ++k) {
        const double z = station_z(k);
        const double c = 1.0 - (1.0 - p.taper) * z / b, xle = z * tsw;
        const double phi_k = phi_b * k / ns;
        const double rho = R * std::cos(phi_k), zq = R * std::sin(phi_k);
        std::vector<V3> P(ni);
        for (uint32_t i = 0; i < ni; ++i) P[i] = {xle + c * ux[i], c * uy[i], z};
        // extrusion directions: section normals smoothed along the ring ([1 2 1] passes), so
        // the layers fan out gradually around the sharp trailing edge
        std::vector<double> nxs(ni), nys(ni);
        for (uint32_t i = 0; i < ni; ++i) {
            const V3 &a = P[(i + 1) % ni], &bb = P[(i + ni - 1) % ni];
            const double tx = a.x - bb.x, ty = a.y - bb.y, tl = std::sqrt(tx * tx + ty * ty);
            nxs[i] = -ty / tl;
            nys[i] = tx / tl;
        }
        for (int pass = 0; pass < 8; ++pass) {
            std::vector<double> ax(ni), ay(ni);
            for (uint32_t i = 0; i < ni; ++i) {
                const uint32_t l = (i + ni - 1) % ni, r = (i + 1) % ni;
                ax[i] = nxs[l] + 2 * nxs[i] + nxs[r];
                ay[i] = nys[l] + 2 * nys[i] + nys[r];
                const double al = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
                ax[i] /= al;
                ay[i] /= al;
            }
            nxs.swap(ax);
            nys.swap(ay);
        }
        for (uint32_t i = 0; i < ni; ++i) {
            // Off the symmetry plane the layers leave the surface along its normal in 3D: the
            // section normal made orthogonal to the spanwise tangent dP/dz (sweep and taper).  A
            // wall cell whose centroid sits off its wall face's normal line lets mass through a
            // slip wall (the ghost takes the cell's unmirrored gradient), which on the swept
            // leading edge fed a growing density spike.
            double nx = nxs[i], ny = nys[i], nz = 0.0;
            if (k > 0) {
                const double tx = tsw - (1.0 - p.taper) / b * ux[i], ty = -(1.0 - p.taper) / b * uy[i], tz = 1.0;
                const double a = (nx * tx + ny * ty) / (tx * tx + ty * ty + tz * tz);
                nx -= a * tx;
                ny -= a * ty;
                nz = -a * tz;
                // no wall-layer node may drop below the mid-height of the previous station
                // (near the root the tilt would take trailing-edge layers under the symmetry plane)
                const double lim = 0.5 * (z - station_z(k - 1)) / d[np] * std::sqrt(nx * nx + ny * ny + nz * nz);
                nz = std::max(-lim, std::min(lim, nz));
                const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
                nx /= l;
                ny /= l;
                nz /= l;
            }
            for (uint32_t j = 0; j <= np; ++j) X[nid(i, j, k)] = {P[i].x + d[j] * nx, P[i].y + d[j] * ny, z + d[j] * nz};
            const V3 Pp = X[nid(i, np, k)];
            const V3 Q = {cx0 + rho * std::cos(th[i]), cy0 + rho * std::sin(th[i]), zq};
            const double L = std::sqrt((Q.x - Pp.x) * (Q.x - Pp.x) + (Q.y - Pp.y) * (Q.y - Pp.y) + (Q.z - Pp.z) * (Q.z - Pp.z));
            require(L > 2 * h_tet, "wing: far field too close to the prism layers");
            const double r = stretch_ratio(h_tet / L, nt);
            std::vector<double> acc(nt + 1);
            {
                double h = h_tet / L;
                acc[0] = 0;
                for (uint32_t j = 1; j <= nt; ++j) {
                    acc[j] = acc[j - 1] + h;
                    h *= r;
                }
            }
            for (uint32_t j = 1; j <= nt; ++j) {
                const double s = j == nt ? 1.0 : acc[j] / acc[nt];
                X[nid(i, np + j, k)] = {Pp.x + s * (Q.x - Pp.x), Pp.y + s * (Q.y - Pp.y), Pp.z + s * (Q.z - Pp.z)};
            }
        }
    }
    // Stations above the tip: the section closes to a plate of zero thickness (its upper and
    // lower ring-0 nodes are merged, like a C-grid's wake cut), so the tip is capped by a low
    // ridge between the tip station and the first station above it, and no cells are needed
    // inside the section.  The stations blend from the tip-station layout to a top layout on
    // the dome: the plate widened about the tip's mid-chord to half the radius of the circle of
    // latitude phi_top (cells above the tip grow with height instead of forming a thin column),
    // rings spread from it to that circle, every node lifted onto the sphere.  Outer-ring nodes
    // move along the sphere at every station.
    {
        const double c_tip = p.taper, xm = b * tsw + 0.5 * c_tip;
        const double rho_top = R * std::cos(phi_top);
        // Top layout in elliptic coordinates about the plate: x = xm + ae cosh(mu) cos(nu),
        // y = -ae sinh(mu) sin(nu), nu = the section parameter of node i.  The map is conformal,
        // so the cells' shape follows the steps in mu against those in nu; the plate's half width
        // ae is at least half the tip chord.  Ring j sits on the ellipse at distance t_j from the
        // plate (ae sinh(mu_j) = t_j), the t_j growing geometrically from 0.07 tip chords to the
        // outermost ellipse: evenly spaced mu would put the first ring 0.015 chords from the
        // plate, under cap layers chords tall.  The outermost ellipse is bent onto the circle of
        // latitude phi_top, the correction growing as cosh(mu) - 1: the rings' spacing next to
        // the plate's ends grows that way, so the correction never turns them back.
        const double ae = std::max(0.5 * c_tip, rho_top / std::cosh(2.0 * pi * nj / ni));
        require(rho_top > 2 * ae, "wing: far field too close to the tip");
        const double mu_max = std::acosh(rho_top / ae);
        std::vector<double> mu_j(nj + 1, 0.0);
        {
            const double tn = ae * std::sinh(mu_max), t1 = std::min(0.07 * c_tip, 0.5 * tn / nj);
            const double g = stretch_ratio(t1 / tn, nj);
            std::vector<double> tt(nj + 1, 0.0);
            double acc = 0, h = t1;
            for (uint32_t j = 1; j <= nj; ++j) {
                acc += h;
                h *= g;
                tt[j] = acc;
            }
            for (uint32_t j = 1; j < nj; ++j) mu_j[j] = std::asinh(tn * (tt[j] / tt[nj]) / ae);
            mu_j[nj] = mu_max;
        }
        std::vector<V3> T((size_t)ni * (nj + 1));
        for (uint32_t i = 0; i < ni; ++i) {
            const double nu = beta_of[i];
            const double ox = cx0 + rho_top * std::cos(th[i]), oy = cy0 + rho_top * std::sin(th[i]);
            const double ex = xm + ae * std::cosh(mu_max) * std::cos(nu), ey = -ae * std::sinh(mu_max) * std::sin(nu);
            for (uint32_t j = 0; j <= nj; ++j) {
                const double mu = mu_j[j], w = (std::cosh(mu) - 1.0) / (std::cosh(mu_max) - 1.0);
                const double x = xm + ae * std::cosh(mu) * std::cos(nu) + w * (ox - ex);
                const double y = (j == 0 ? 0.0 : -ae * std::sinh(mu) * std::sin(nu)) + w * (oy - ey);
                const double rr = (x - cx0) * (x - cx0) + (y - cy0) * (y - cy0);
                T[(size_t)j * ni + i] = {x, y, j == nj ? R * std::sin(phi_top) : std::sqrt(std::max(0.0, R * R - rr))};
            }
        }
        // Cap heights: steps growing by 1.6 from the last span step up to a largest step hm,
        // which the remaining steps keep, hm chosen so the steps reach the dome (a plain
        // geometric progression over n_cap steps ends in one several times the mean, and the
        // column's top cells with it).  Fractions sq[q] of the height H from the tip to the dome.
        const double dz_tip = station_z(ns) - station_z(ns - 1);
        const double H = R - b;
        std::vector<double> sq(nc + 1, 0.0);
        {
            const double g = 1.6;
            auto steps = [&](double hm, bool store) {
                double acc = 0, h = dz_tip;
                for (uint32_t q = 1; q <= nc; ++q) {
                    acc += std::min(h, hm);
                    h *= g;
                    if (store) sq[q] = acc;
                }
                return acc;
            };
            double hm = H;   // too few stations to reach the dome with steps growing by g: no cap
            if (steps(H, false) > H) {
                double lo = dz_tip, hi = H;
                for (int it = 0; it < 200; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    (steps(mid, false) > H ? hi : lo) = mid;
                }
                hm = 0.5 * (lo + hi);
            }
            steps(hm, true);
            for (uint32_t q = 1; q < nc; ++q) sq[q] /= sq[nc];
            sq[nc] = 1.0;
        }
        while (n_taper + 1 <= nc / 2 && sq[n_taper + 1] * H <= 2.0 * p.thickness * p.taper) ++n_taper;
        // Station q above the tip.  Ring j's in-plane layout is B + s (T - B), B the tip station,
        // T the top layout, with s moving from the geometric cap fraction sq[q] at the plate (in
        // step with the heights, so no station slides over the one below) to s_out(q) at the
        // far-field ring, where s_out puts that ring (the blend of B's and T's
        // circles) on the circle of latitude phi_b + (phi_top - phi_b) q / n_cap.  The weight
        // w_j is ring j's mean fraction of the path from the section to the far field on the tip
        // station, the same for every node of a ring, so every ring is a convex blend of two
        // valid layouts with one weight and the lines across rings do not fold.  Heights rise
        // from the tip station's by the geometric cap increment at the plate and by the
        // latitude's increment at the far field, blended with w_j.  The top station keeps T, on
        // the sphere.
        std::vector<double> wj(nj + 1, 0.0);
        for (uint32_t i = 0; i < ni; ++i) {
            double acc = 0;
            std::vector<double> l(nj + 1, 0.0);
            for (uint32_t j = 1; j <= nj; ++j) {
                const V3 a0 = X[nid(i, j - 1, ns)], a1 = X[nid(i, j, ns)];
                acc += std::hypot(a1.x - a0.x, a1.y - a0.y);
                l[j] = acc;
            }
            for (uint32_t j = 1; j <= nj; ++j) wj[j] += l[j] / acc / ni;
        }
        wj[nj] = 1.0;
        const double rho_b = R * std::cos(phi_b);
        for (uint32_t q = 1; q < nc; ++q) {
            const double phi = phi_b + (phi_top - phi_b) * q / nc;
            const double s_out = (rho_b - R * std::cos(phi)) / (rho_b - rho_top), fz = R * std::sin(phi);
            // the section's thickness tapers to the plate over the first n_taper stations
            const double taper = q < n_taper ? 1.0 - (double)q / n_taper : 0.0;
            for (uint32_t i = 0; i < ni; ++i) {
                // on the plate, upper node i coincides with lower node ni - i
                const uint32_t il = 2 * i > ni && q >= n_taper ? ni - i : i;
                const V3 Bl = X[nid(il, 0, ns)], Dl = T[il];
                const double zp = Bl.z + sq[q] * (Dl.z - Bl.z);
                // the section's surface moves towards the plate by dy0; the rings next to it move
                // with it (fading with w_j), so the wall layers keep their thickness and stay
                // normal to the tapering surface
                const double dy0 = (taper - 1.0) * X[nid(i, 0, ns)].y;
                for (uint32_t j = 0; j <= nj; ++j) {
                    const uint32_t ii = j == 0 ? il : i;
                    const V3 B = X[nid(ii, j, ns)], D = T[(size_t)j * ni + ii];
                    const double w = wj[j], sj = (1.0 - w) * sq[q] + w * s_out;
                    const double y = B.y + sj * (D.y - B.y) + (1.0 - sj) * (1.0 - w) * (j == 0 ? (taper - 1.0) * B.y : dy0);
                    // height: the tip station's plus an increment blended from the plate's to
                    // the far-field ring's, but not below the plate's height: the stations are
                    // level around the plate, where the cells are tall and narrow (a sloping
                    // station there feeds the vertical gradient from the in-plane neighbours).
                    // Both terms rise with q, so the stations never cross.
                    const double dz = (1.0 - w) * (zp - Bl.z) + w * (fz - R * std::sin(phi_b));
                    X[nid(i, j, ns + q)] = {B.x + sj * (D.x - B.x), y, std::max(zp, B.z + dz)};
                }
            }
        }
        for (uint32_t j = 0; j <= nj; ++j)
            for (uint32_t i = 0; i < ni; ++i) X[nid(i, j, K)] = T[(size_t)j * ni + i];
    }
    // node ids: upper plate nodes (j = 0, k >= n_span + n_taper, i > ni/2) are the lower ones
    std::vector<uint32_t> remap(n_grid), canon;
    canon.reserve(n_grid);
    e = gr_elements();
    e.xyz.reserve(3 * n_grid);
    for (uint32_t k = 0; k <= K; ++k)
        for (uint32_t j = 0; j <= nj; ++j)
            for (uint32_t i = 0; i < ni; ++i) {
                const uint32_t g = nid(i, j, k);
                if (j == 0 && k >= ns + n_taper && 2 * i > ni) {
                    remap[g] = remap[nid(ni - i, 0, k)];
                    continue;
                }
                remap[g] = (uint32_t)(e.xyz.size() / 3);
                canon.push_back(g);
                e.xyz.push_back(X[g].x * p.scale);
                e.xyz.push_back(X[g].y * p.scale);
                e.xyz.push_back(X[g].z * p.scale);
            }
    auto pos = [&](uint32_t n) { return V3{e.xyz[3 * n], e.xyz[3 * n + 1], e.xyz[3 * n + 2]}; };
    e.starts.push_back(0);
    auto add = [&](uint32_t kind, const uint32_t *v, int cnt) {
        e.kinds.push_back(kind);
        e.nodes.insert(e.nodes.end(), v, v + cnt);
        e.starts.push_back((uint32_t)e.nodes.size());
    };
    // tetrahedron with positive orientation ((1-0) x (2-0) . (3-0) > 0)
    auto add_tet = [&](uint32_t a, uint32_t bb, uint32_t c, uint32_t dd) {
        const V3 A = pos(a);
        if (det3(sub(pos(bb), A), sub(pos(c), A), sub(pos(dd), A)) < 0) std::swap(bb, c);
        const uint32_t t[4] = {a, bb, c, dd};
        add(GR_TET, t, 4);
    };
    // prism: triangle (a,b,c) at j, (d,e,f) above at j+1
    auto add_prism = [&](uint32_t a, uint32_t bb, uint32_t c, uint32_t dd, uint32_t ee, uint32_t f) {
        const V3 A = pos(a);
        if (det3(sub(pos(bb), A), sub(pos(c), A), sub(pos(dd), A)) < 0) {
            std::swap(bb, c);
            std::swap(ee, f);
        }
        const uint32_t w[6] = {a, bb, c, dd, ee, f};
        add(GR_WEDGE, w, 6);
    };
    // pyramid: base quad (a,b,c,d) in cyclic order, apex e on the positive side
    auto add_pyramid = [&](uint32_t a, uint32_t bb, uint32_t c, uint32_t dd, uint32_t ee) {
        const V3 A = pos(a);
        if (det3(sub(pos(bb), A), sub(pos(dd), A), sub(pos(ee), A)) < 0) std::swap(bb, dd);
        const uint32_t w[5] = {a, bb, c, dd, ee};
        add(GR_PYRAMID, w, 5);
    };
    // Cells of hex (i, j, k), corner[a + 2b + 4c] at offsets a, b, c along i, j, k.  A stretched
    // triangle, in a prism or a tetrahedron, is what makes the unlimited second-order update of
    // large cells unstable; a long prism over a compact triangle is not.  A thin tetrahedron on
    // a wall is bad too: its centroid sits far off its wall face's normal line, which lets mass
    // through a slip wall.
    //  - Wall layers (j < n_prism) up to the top of the tip's taper (k < n_span + n_taper):
    //    prisms extruded along j, triangles on constant-j faces split along (i, k) -> (i+1, k+1).
    //  - Outside them below the tip (k < n_span): the 6-tetrahedron split about the
    //    corner-0 -> corner-7 diagonal, whose face diagonals match the prisms'.
    //  - Above the tip (k >= n_span, j > n_prism): prisms extruded along k (upwards), triangles
    //    on constant-k faces split along (i, j) -> (i+1, j+1) as the tetrahedra below; cells
    //    there are tall near the plate.  Their first layer (j = n_prism) up to the transition row
    //    is a pyramid on the quad of the k-extruded prism outside it plus the four tetrahedra of
    //    the split not touching that quad.
    //  - The transition row k = n_span + n_taper (on the plate, no wall): wall-layer hexes as a
    //    pyramid on the j-extruded prism's quad below plus four tetrahedra; above it every cell
    //    is a k-extruded prism.
    // A slice i built mirrored in i flips every diagonal except those on constant-i faces, so it
    // still meets its neighbours.  The plate's lower column i and upper column n_around-1-i
    // share their face on the transition row, so exactly one of each pair is mirrored: the
    // upper half except the leading-edge slice n_around/2, plus the lower slice n_around/2 - 1
    // -- the slices next to either end of the plate (n_around/2 - 1 and n_around - 1) must be
    // mirrored, since there the unmirrored triangle (i, 0), (i+1, 0), (i+1, 1) of a k-extruded
    // prism has its corners on the plate's line.
    require(n_taper < nc, "wing: need more stations above the tip");
    const uint32_t k_tr = ns + n_taper;
    for (uint32_t k = 0; k < K; ++k)
        for (uint32_t j = 0; j < nj; ++j)
            for (uint32_t i = 0; i < ni; ++i) {
                const int mirror = ((2 * i >= ni && 2 * i != ni) || i + 1 == ni / 2) ? 1 : 0;
                uint32_t cr[8];
                for (int o = 0; o < 8; ++o) {
                    const int m = o ^ mirror;
                    cr[o] = remap[nid(i + (m & 1), j + ((m >> 1) & 1), k + ((m >> 2) & 1))];
                }
                if (j < np && k < k_tr) {
                    add_prism(cr[0], cr[1], cr[5], cr[2], cr[3], cr[7]);
                    add_prism(cr[0], cr[5], cr[4], cr[2], cr[7], cr[6]);
                } else if (j < np && k == k_tr) {
                    add_pyramid(cr[0], cr[1], cr[3], cr[2], cr[7]);
                    add_tet(cr[0], cr[1], cr[5], cr[7]);
                    add_tet(cr[0], cr[4], cr[5], cr[7]);
                    add_tet(cr[0], cr[2], cr[6], cr[7]);
                    add_tet(cr[0], cr[4], cr[6], cr[7]);
                } else if (j == np && k >= ns && k <= k_tr) {
                    add_pyramid(cr[2], cr[3], cr[7], cr[6], cr[0]);
                    add_tet(cr[0], cr[1], cr[3], cr[7]);
                    add_tet(cr[0], cr[1], cr[5], cr[7]);
                    add_tet(cr[0], cr[4], cr[5], cr[7]);
                    add_tet(cr[0], cr[4], cr[6], cr[7]);
                } else if (k >= ns) {
                    add_prism(cr[0], cr[1], cr[3], cr[4], cr[5], cr[7]);
                    add_prism(cr[0], cr[3], cr[2], cr[4], cr[7], cr[6]);
                } else {
                    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                    for (const auto &pm : perms) {
                        const int o1 = 1 << pm[0], o2 = o1 | (1 << pm[1]);
                        add_tet(cr[0], cr[o1], cr[o2], cr[7]);
                    }
                }
            }
    auto tag_j = [&](uint32_t n) -> int64_t { return (int64_t)((canon[n] / ni) % (nj + 1)); };
    auto tag_k = [&](uint32_t n) -> int64_t { return (int64_t)(canon[n] / (ni * (nj + 1))); };
    add_boundary_markers(e, {"wall", "farfield", "symmetry"}, [&](const uint32_t *fn, int c) {
        auto all = [&](auto pred) {
            for (int q = 0; q < c; ++q)
                if (!pred(fn[q])) return false;
            return true;
        };
        if (all([&](uint32_t n) { return tag_k(n) == 0; })) return 2;
        if (all([&](uint32_t n) { return tag_j(n) == 0; })) return 0;
        if (all([&](uint32_t n) { return tag_j(n) == (int64_t)nj || tag_k(n) == (int64_t)K; })) return 1;
        return -1;
    });
}

// ---- su2 text format (src/meshio.cu:771-897) -----------------------------------------------

void read_su2(const char *path, gr_elements &e) {
    std::ifstream in(path);
    require(in.good(), std::string("cannot open su2 file ") + path, GR_ERR_IO);
    e = gr_elements();
    std::string tok;
    int dim = 3;
    auto read_elems = [&](uint32_t count, bool with_id, std::vector<uint32_t> &kinds,
                          std::vector<uint32_t> &starts, std::vector<uint32_t> &nodes) {
        starts.assign(1, 0);
        for (uint32_t i = 0; i < count; ++i) {
            uint32_t kind;
            require((bool)(in >> kind), "su2: truncated element list", GR_ERR_IO);
            ElemKind k;
            require(elem_kind(kind, k), "su2: unknown element kind " + std::to_string(kind), GR_ERR_IO);
            kinds.push_back(kind);
            for (int j = 0; j < k.n_nodes; ++j) {
                uint32_t n;
                require((bool)(in >> n), "su2: truncated element", GR_ERR_IO);
                nodes.push_back(n);
            }
            starts.push_back((uint32_t)nodes.size());
            if (with_id) {
                uint64_t id;
                require((bool)(in >> id) && id == i, "su2: element id out of sequence", GR_ERR_IO);
            }
        }
    };
    while (in >> tok) {
        if (tok == "NDIME=") {
            require((bool)(in >> dim) && dim == 3, "su2: mesh must be 3D", GR_ERR_IO);
        } else if (tok == "NPOIN=") {
            uint64_t n;
            require((bool)(in >> n), "su2: bad NPOIN", GR_ERR_IO);
            // NPOIN= may carry a second count on the same line (gmsh writes one)
            std::string rest;
            std::getline(in, rest);
            e.xyz.resize(3 * n);
            for (uint64_t i = 0; i < n; ++i) {
                std::string line;
                require((bool)std::getline(in, line), "su2: truncated point list", GR_ERR_IO);
                std::istringstream ls(line);
                double x = 0, y = 0, z = 0;
                require((bool)(ls >> x >> y >> z), "su2: bad point line", GR_ERR_IO);
                e.xyz[3 * i] = x;
                e.xyz[3 * i + 1] = y;
                e.xyz[3 * i + 2] = z;
            }
        } else if (tok == "NELEM=") {
            uint32_t n;
            require((bool)(in >> n), "su2: bad NELEM", GR_ERR_IO);
            read_elems(n, true, e.kinds, e.starts, e.nodes);
        } else if (tok == "NMARK=") {
            uint32_t nm;
            require((bool)(in >> nm), "su2: bad NMARK", GR_ERR_IO);
            for (uint32_t m = 0; m < nm; ++m) {
                std::string t1, name, t2;
                uint32_t cnt;
                require((bool)(in >> t1 >> name) && t1 == "MARKER_TAG=", "su2: expected MARKER_TAG=", GR_ERR_IO);
                require((bool)(in >> t2 >> cnt) && t2 == "MARKER_ELEMS=", "su2: expected MARKER_ELEMS=", GR_ERR_IO);
                Marker mk;
                mk.name = name;
                read_elems(cnt, false, mk.kinds, mk.starts, mk.nodes);
                e.markers.push_back(std::move(mk));
            }
        }
    }
    require(!e.kinds.empty(), "su2: no elements", GR_ERR_IO);
    check_elements(e.n_nodes(), e.n_elems(), e.kinds.data(), e.starts.data(), e.nodes.data(), false);
    for (const auto &m : e.markers)
        check_elements(e.n_nodes(), (uint32_t)m.kinds.size(), m.kinds.data(), m.starts.data(),
                       m.nodes.data(), true);
}

void write_su2(const gr_elements &e, const char *path) {
    FILE *f = std::fopen(path, "w");
    require(f != nullptr, std::string("cannot open for writing ") + path, GR_ERR_IO);
    std::fprintf(f, "NDIME= 3\nNELEM= %u\n", e.n_elems());
    for (uint32_t i = 0; i < e.n_elems(); ++i) {
        std::fprintf(f, "%u", e.kinds[i]);
        for (uint32_t q = e.starts[i]; q < e.starts[i + 1]; ++q) std::fprintf(f, " %u", e.nodes[q]);
        std::fprintf(f, " %u\n", i);
    }
    std::fprintf(f, "NPOIN= %u\n", e.n_nodes());
    for (uint32_t i = 0; i < e.n_nodes(); ++i)
        std::fprintf(f, "%.17g %.17g %.17g %u\n", e.xyz[3 * i], e.xyz[3 * i + 1], e.xyz[3 * i + 2], i);
    std::fprintf(f, "NMARK= %u\n", (uint32_t)e.markers.size());
    for (const auto &m : e.markers) {
        std::fprintf(f, "MARKER_TAG= %s\nMARKER_ELEMS= %u\n", m.name.c_str(), (uint32_t)m.kinds.size());
        for (size_t i = 0; i < m.kinds.size(); ++i) {
            std::fprintf(f, "%u", m.kinds[i]);
            for (uint32_t q = m.starts[i]; q < m.starts[i + 1]; ++q) std::fprintf(f, " %u", m.nodes[q]);
            std::fprintf(f, "\n");
        }
    }
    bool ok = std::ferror(f) == 0;
    ok = (std::fclose(f) == 0) && ok;
    require(ok, std::string("write failed: ") + path, GR_ERR_IO);
}

}  // namespace gr

// ---- C ABI -------------------------------------------------------------------------------

using gr::guarded;
using gr::require;

extern "C" {

int gr_elements_create(uint32_t n_nodes, const double *xyz, uint32_t n_elems,
                       const uint32_t *kinds, const uint32_t *starts, const uint32_t *nodes,
                       gr_elements **out) {
    return guarded([&] {
        require(out && xyz && kinds && starts && nodes, "null argument");
        require(n_elems > 0, "no elements");
        gr::check_elements(n_nodes, n_elems, kinds, starts, nodes, false);
        auto *e = new gr_elements();
        e->xyz.assign(xyz, xyz + 3 * (size_t)n_nodes);
        e->kinds.assign(kinds, kinds + n_elems);
        e->starts.assign(starts, starts + n_elems + 1);
        e->nodes.assign(nodes, nodes + starts[n_elems]);
        *out = e;
    });
}

int gr_elements_add_marker(gr_elements *e, const char *name, uint32_t n, const uint32_t *kinds,
                           const uint32_t *starts, const uint32_t *nodes) {
    return guarded([&] {
        require(e && name && kinds && starts && nodes, "null argument");
        require(std::strlen(name) > 0 && std::strlen(name) < 128, "marker name must have 1..127 chars");
        gr::check_elements(e->n_nodes(), n, kinds, starts, nodes, true);
        gr::Marker m;
        m.name = name;
        m.kinds.assign(kinds, kinds + n);
        m.starts.assign(starts, starts + n + 1);
        m.nodes.assign(nodes, nodes + starts[n]);
        e->markers.push_back(std::move(m));
    });
}

int gr_elements_read_su2(const char *path, gr_elements **out) {
    return guarded([&] {
        require(path && out, "null argument");
        auto *e = new gr_elements();
        try {
            gr::read_su2(path, *e);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int gr_elements_write_su2(const gr_elements *e, const char *path) {
    return guarded([&] {
        require(e && path, "null argument");
        gr::write_su2(*e, path);
    });
}

int gr_elements_info_get(const gr_elements *e, gr_elements_info *out) {
    return guarded([&] {
        require(e && out, "null argument");
        out->n_nodes = e->n_nodes();
        out->n_elems = e->n_elems();
        out->n_elem_nodes = (uint32_t)e->nodes.size();
        out->n_markers = (uint32_t)e->markers.size();
    });
}

int gr_elements_get(const gr_elements *e, double *xyz, uint32_t *kinds, uint32_t *starts,
                    uint32_t *nodes) {
    return guarded([&] {
        require(e != nullptr, "null argument");
        if (xyz) std::copy(e->xyz.begin(), e->xyz.end(), xyz);
        if (kinds) std::copy(e->kinds.begin(), e->kinds.end(), kinds);
        if (starts) std::copy(e->starts.begin(), e->starts.end(), starts);
        if (nodes) std::copy(e->nodes.begin(), e->nodes.end(), nodes);
    });
}

int gr_elements_get_marker(const gr_elements *e, uint32_t m, char name[128], uint32_t *n,
                           uint32_t *n_nodes, uint32_t *kinds, uint32_t *starts, uint32_t *nodes) {
    return guarded([&] {
        require(e != nullptr, "null argument");
        require(m < e->markers.size(), "marker index out of range");
        const auto &mk = e->markers[m];
        if (name) {
            std::memset(name, 0, 128);
            std::strncpy(name, mk.name.c_str(), 127);
        }
        if (n) *n = (uint32_t)mk.kinds.size();
        if (n_nodes) *n_nodes = (uint32_t)mk.nodes.size();
        if (kinds) std::copy(mk.kinds.begin(), mk.kinds.end(), kinds);
        if (starts) std::copy(mk.starts.begin(), mk.starts.end(), starts);
        if (nodes) std::copy(mk.nodes.begin(), mk.nodes.end(), nodes);
    });
}

void gr_elements_free(gr_elements *e) { delete e; }

int gr_gen_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, int tets,
               double jitter, uint64_t seed, gr_elements **out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        auto *e = new gr_elements();
        try {
            gr::gen_box(nx, ny, nz, lx, ly, lz, tets, jitter, seed, *e);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int gr_gen_ogrid(uint32_t n_theta, uint32_t n_r, uint32_t n_z, double r0, double r1,
                 double ratio, double dz, gr_elements **out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        auto *e = new gr_elements();
        try {
            gr::gen_ogrid(n_theta, n_r, n_z, r0, r1, ratio, dz, *e);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int gr_gen_wing(const gr_wing_params *p, gr_elements **out) {
    return guarded([&] {
        require(p && out, "null argument");
        auto *e = new gr_elements();
        try {
            gr::gen_wing(*p, *e);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int gr_gen_cgrid_naca(uint32_t n_i, uint32_t n_j, uint32_t n_wake, double thickness, double radius,
                      double first_height, double dz, gr_elements **out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        auto *e = new gr_elements();
        try {
            gr::gen_cgrid(n_i, n_j, n_wake, thickness, radius, first_height, dz, *e);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

}  // extern "C"
