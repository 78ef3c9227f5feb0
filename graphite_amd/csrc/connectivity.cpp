// Host connectivity: unique faces and CSR arrays in cuMeshFromSu2's numbering, .fvm I/O,
// the 1-D channel mesh, RCM renumbering and face colouring.
//
// cuMeshFromSu2 (src/meshio.cu:713-1138) numbers faces by walking a chained hash table of
// n_elems+1 buckets: buckets in index order, each chain in creation order.  A face's bucket
// is ((prod_k (998244353 + 2*s_k)) mod 2^32) / 2 mod (n_elems+1) over its sorted node ids s,
// and its place in the chain is the order in which faces were first met (elements in order,
// each element's local faces in table order).  This builder reaches the same numbering
// without the table: it sorts the element faces by key to find each face's owners and its
// first occurrence, then sorts the unique faces by (bucket, first occurrence).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <queue>

#include "gr_internal.hpp"

namespace gr {

static inline uint32_t face_bucket(const uint32_t *sorted, int n, uint32_t table_size) {
    uint32_t v = 1u;
    for (int i = 0; i < n; ++i) v *= 998244353u + 2u * sorted[i];
    return (v / 2u) % table_size;
}

void HostMesh::validate() const {
    const uint64_t NG = (uint64_t)n_cells + n_ghosts;
    require(nodes.size() == 3ull * n_nodes, "mesh: node array size mismatch", GR_ERR_TOPOLOGY);
    require(face_nodes_starts.size() == (size_t)n_faces + 1 && face_nodes_starts[0] == 0,
            "mesh: face starts size mismatch", GR_ERR_TOPOLOGY);
    require(face_nodes.size() == face_nodes_starts[n_faces], "mesh: face node count mismatch", GR_ERR_TOPOLOGY);
    for (uint32_t f = 0; f < n_faces; ++f) {
        require(face_nodes_starts[f + 1] >= face_nodes_starts[f] + 3, "mesh: face with fewer than 3 nodes",
                GR_ERR_TOPOLOGY);
    }
    for (uint32_t v : face_nodes) require(v < n_nodes, "mesh: face node index out of range", GR_ERR_TOPOLOGY);
    require(cell_faces_starts.size() == NG + 1 && cell_faces_starts[0] == 0, "mesh: cell starts size mismatch",
            GR_ERR_TOPOLOGY);
    const size_t slots = cell_faces_starts[NG];
    require(cell_faces.size() == slots && cell_connects.size() == slots, "mesh: cell slot count mismatch",
            GR_ERR_TOPOLOGY);
    for (uint64_t c = 0; c < NG; ++c) {
        uint32_t s = cell_faces_starts[c], e = cell_faces_starts[c + 1];
        require(e >= s, "mesh: decreasing cell starts", GR_ERR_TOPOLOGY);
        if (c < n_cells) require(e - s >= 2 && e - s <= 64, "mesh: interior cell with <2 or >64 faces", GR_ERR_TOPOLOGY);
        else require(e - s == 1, "mesh: ghost cell must have exactly one face", GR_ERR_TOPOLOGY);
        for (uint32_t q = s; q < e; ++q) {
            require(cell_faces[q] < n_faces, "mesh: cell face index out of range", GR_ERR_TOPOLOGY);
            require(cell_connects[q] < NG && cell_connects[q] != c, "mesh: cell neighbour index out of range",
                    GR_ERR_TOPOLOGY);
            if (c >= n_cells)
                require(cell_connects[q] < n_cells, "mesh: ghost cell must connect to an interior cell",
                        GR_ERR_TOPOLOGY);
        }
    }
    require(ghost_cell_ids.size() == n_ghosts && ghost_cell_starts.size() == (size_t)n_regions + 1,
            "mesh: region array size mismatch", GR_ERR_TOPOLOGY);
    require(ghost_cell_starts[0] == 0 && ghost_cell_starts[n_regions] == n_ghosts, "mesh: bad region starts",
            GR_ERR_TOPOLOGY);
    for (uint32_t r = 0; r < n_regions; ++r)
        require(ghost_cell_starts[r + 1] >= ghost_cell_starts[r], "mesh: decreasing region starts", GR_ERR_TOPOLOGY);
    for (uint32_t g = 0; g < n_ghosts; ++g)
        require(ghost_cell_ids[g] == n_cells + g,
                "mesh: ghost cells must be numbered n_cells.. in region order (as cuMeshFromSu2 writes them)",
                GR_ERR_TOPOLOGY);
    require(region_names.size() == n_regions, "mesh: region name count mismatch", GR_ERR_TOPOLOGY);
}

HostMesh build_mesh(const gr_elements &e) {
    const uint32_t NE = e.n_elems();
    require(NE > 0, "mesh: no elements");
    uint32_t n_ghosts = 0;
    for (const auto &m : e.markers) n_ghosts += (uint32_t)m.kinds.size();

    // 1. element faces in insertion order t (elements in order, local faces in table order)
    std::vector<uint32_t> elem_face_off(NE + 1, 0);
    for (uint32_t el = 0; el < NE; ++el) {
        ElemKind k;
        require(elem_kind(e.kinds[el], k), "mesh: unknown element kind", GR_ERR_TOPOLOGY);
        elem_face_off[el + 1] = elem_face_off[el] + (uint32_t)k.n_faces;
    }
    const uint32_t T = elem_face_off[NE];
    struct Key { uint32_t s[4]; uint32_t n; };
    std::vector<Key> key(T);
    std::vector<uint32_t> t_elem(T);
    std::vector<uint8_t> t_local(T);
    for (uint32_t el = 0; el < NE; ++el) {
        ElemKind k;
        elem_kind(e.kinds[el], k);
        for (int f = 0; f < k.n_faces; ++f) {
            uint32_t t = elem_face_off[el] + f;
            Key &K = key[t];
            K.n = 0;
            for (int j = 0; j < 4; ++j)
                if (k.f[f][j] >= 0) K.s[K.n++] = e.nodes[e.starts[el] + k.f[f][j]];
            for (uint32_t j = K.n; j < 4; ++j) K.s[j] = 0xFFFFFFFFu;
            std::sort(K.s, K.s + K.n);
            t_elem[t] = el;
            t_local[t] = (uint8_t)f;
        }
    }
    auto key_less = [&](uint32_t a, uint32_t b) {
        const Key &A = key[a], &B = key[b];
        if (A.n != B.n) return A.n < B.n;
        for (int j = 0; j < 4; ++j)
            if (A.s[j] != B.s[j]) return A.s[j] < B.s[j];
        return a < b;   // stable: first occurrence first
    };
    auto key_eq = [&](uint32_t a, uint32_t b) {
        const Key &A = key[a], &B = key[b];
        return A.n == B.n && std::memcmp(A.s, B.s, sizeof A.s) == 0;
    };
    std::vector<uint32_t> order(T);
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), key_less);

    // 2. unique faces: first occurrence and owners (elements in insertion order)
    std::vector<uint32_t> uface_of_t(T);
    std::vector<uint32_t> u_first;          // t of the first occurrence
    std::vector<uint32_t> u_owner0, u_owner1, u_count;
    for (uint32_t a = 0; a < T;) {
        uint32_t b = a + 1;
        while (b < T && key_eq(order[a], order[b])) ++b;
        uint32_t u = (uint32_t)u_first.size();
        u_first.push_back(order[a]);
        u_owner0.push_back(t_elem[order[a]]);
        u_owner1.push_back(b - a > 1 ? t_elem[order[a + 1]] : 0xFFFFFFFFu);
        u_count.push_back(b - a);
        for (uint32_t q = a; q < b; ++q) uface_of_t[order[q]] = u;
        a = b;
    }
    const uint32_t U = (uint32_t)u_first.size();

    // 3. ghosts: n_elems + running index in marker order, attached to their face
    std::vector<uint32_t> ghost_face(n_ghosts);
    {
        uint32_t g = 0;
        for (const auto &mk : e.markers) {
            for (size_t i = 0; i < mk.kinds.size(); ++i, ++g) {
                Key K;
                K.n = mk.starts[i + 1] - mk.starts[i];
                require(K.n >= 3 && K.n <= 4, "mesh: marker element must be a triangle or quad", GR_ERR_TOPOLOGY);
                for (uint32_t j = 0; j < K.n; ++j) K.s[j] = mk.nodes[mk.starts[i] + j];
                for (uint32_t j = K.n; j < 4; ++j) K.s[j] = 0xFFFFFFFFu;
                std::sort(K.s, K.s + K.n);
                // binary search among the sorted element faces
                auto it = std::lower_bound(order.begin(), order.end(), K, [&](uint32_t t, const Key &k2) {
                    const Key &A = key[t];
                    if (A.n != k2.n) return A.n < k2.n;
                    for (int j = 0; j < 4; ++j)
                        if (A.s[j] != k2.s[j]) return A.s[j] < k2.s[j];
                    return false;
                });
                bool found = it != order.end() && key[*it].n == K.n && std::memcmp(key[*it].s, K.s, sizeof K.s) == 0;
                require(found, "mesh: marker '" + mk.name + "' element " + std::to_string(i) +
                                   " is not a face of the volume mesh", GR_ERR_TOPOLOGY);
                uint32_t u = uface_of_t[*it];
                if (u_count[u] == 1) u_owner1[u] = NE + g;
                u_count[u]++;
                ghost_face[g] = u;
            }
        }
    }
    for (uint32_t u = 0; u < U; ++u) {
        if (u_count[u] != 2) {
            const Key &K = key[u_first[u]];
            std::string nodes;
            for (uint32_t j = 0; j < K.n; ++j) nodes += (j ? "," : "") + std::to_string(K.s[j]);
            throw Error(GR_ERR_TOPOLOGY, "mesh: face (" + nodes + ") has " + std::to_string(u_count[u]) +
                                             " cells, must be 2 (boundary faces need a marker)");
        }
    }

    // 4. face ids: bucket order, then order of first occurrence within a bucket
    const uint32_t table_size = NE + 1;
    std::vector<uint32_t> u_bucket(U);
    for (uint32_t u = 0; u < U; ++u) u_bucket[u] = face_bucket(key[u_first[u]].s, (int)key[u_first[u]].n, table_size);
    std::vector<uint32_t> fid_order(U);
    std::iota(fid_order.begin(), fid_order.end(), 0u);
    std::sort(fid_order.begin(), fid_order.end(), [&](uint32_t a, uint32_t b) {
        if (u_bucket[a] != u_bucket[b]) return u_bucket[a] < u_bucket[b];
        return u_first[a] < u_first[b];
    });
    std::vector<uint32_t> fid_of_u(U);
    for (uint32_t f = 0; f < U; ++f) fid_of_u[fid_order[f]] = f;

    HostMesh m;
    m.n_nodes = e.n_nodes();
    m.n_cells = NE;
    m.n_ghosts = n_ghosts;
    m.n_regions = (uint32_t)e.markers.size();
    m.n_faces = U;
    m.nodes = e.xyz;
    // face nodes in the local order of the first element that met the face
    m.face_nodes_starts.resize(U + 1);
    m.face_nodes_starts[0] = 0;
    for (uint32_t f = 0; f < U; ++f) m.face_nodes_starts[f + 1] = m.face_nodes_starts[f] + key[u_first[fid_order[f]]].n;
    m.face_nodes.resize(m.face_nodes_starts[U]);
    for (uint32_t f = 0; f < U; ++f) {
        uint32_t t = u_first[fid_order[f]];
        uint32_t el = t_elem[t];
        ElemKind k;
        elem_kind(e.kinds[el], k);
        uint32_t off = m.face_nodes_starts[f];
        for (int j = 0; j < 4; ++j)
            if (k.f[t_local[t]][j] >= 0) m.face_nodes[off++] = e.nodes[e.starts[el] + k.f[t_local[t]][j]];
    }
    // cell CSR: element-local face order; ghosts have one face
    const uint32_t NG = NE + n_ghosts;
    m.cell_faces_starts.resize(NG + 1);
    m.cell_faces_starts[0] = 0;
    for (uint32_t c = 0; c < NE; ++c) m.cell_faces_starts[c + 1] = m.cell_faces_starts[c] + (elem_face_off[c + 1] - elem_face_off[c]);
    for (uint32_t g = 0; g < n_ghosts; ++g) m.cell_faces_starts[NE + g + 1] = m.cell_faces_starts[NE + g] + 1;
    m.cell_faces.resize(m.cell_faces_starts[NG]);
    m.cell_connects.resize(m.cell_faces_starts[NG]);
    for (uint32_t c = 0; c < NE; ++c) {
        for (uint32_t t = elem_face_off[c]; t < elem_face_off[c + 1]; ++t) {
            uint32_t u = uface_of_t[t];
            uint32_t slot = m.cell_faces_starts[c] + (t - elem_face_off[c]);
            m.cell_faces[slot] = fid_of_u[u];
            m.cell_connects[slot] = (u_owner0[u] == c) ? u_owner1[u] : u_owner0[u];
        }
    }
    for (uint32_t g = 0; g < n_ghosts; ++g) {
        uint32_t u = ghost_face[g];
        uint32_t slot = m.cell_faces_starts[NE + g];
        m.cell_faces[slot] = fid_of_u[u];
        m.cell_connects[slot] = (u_owner0[u] == NE + g) ? u_owner1[u] : u_owner0[u];
    }
    m.ghost_cell_ids.resize(n_ghosts);
    for (uint32_t g = 0; g < n_ghosts; ++g) m.ghost_cell_ids[g] = NE + g;
    m.ghost_cell_starts.resize(m.n_regions + 1);
    m.ghost_cell_starts[0] = 0;
    for (uint32_t r = 0; r < m.n_regions; ++r)
        m.ghost_cell_starts[r + 1] = m.ghost_cell_starts[r] + (uint32_t)e.markers[r].kinds.size();
    for (const auto &mk : e.markers) m.region_names.push_back(mk.name);
    m.validate();
    return m;
}

HostMesh make_1d(size_t n, double x0, double x1) {
    // cuMeshMake1D, src/mesh.cu:540-634 (which places nodes at i/n whatever x0, x1 are).
    require(n >= 2 && n < 0x3FFFFFF0ull, "1d mesh: need 2 <= n < 2^30");
    HostMesh m;
    const uint32_t N = (uint32_t)n;
    m.n_cells = N;
    m.n_faces = N + 1;
    m.n_nodes = 4 * (N + 1);
    m.n_ghosts = 2;
    m.n_regions = 2;
    m.nodes.resize(3 * (size_t)m.n_nodes);
    for (uint32_t i = 0; i <= N; ++i) {
        double x = x0 + (x1 - x0) * ((double)i / (double)N);
        const double yz[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        for (int k = 0; k < 4; ++k) {
            m.nodes[3 * (4 * i + k)] = x;
            m.nodes[3 * (4 * i + k) + 1] = yz[k][0];
            m.nodes[3 * (4 * i + k) + 2] = yz[k][1];
        }
    }
    m.face_nodes_starts.resize(N + 2);
    m.face_nodes.resize(4 * (N + 1));
    for (uint32_t i = 0; i <= N; ++i) {
        m.face_nodes_starts[i] = 4 * i;
        for (int k = 0; k < 4; ++k) m.face_nodes[4 * i + k] = 4 * i + k;
    }
    m.face_nodes_starts[N + 1] = 4 * (N + 1);
    m.cell_faces_starts.resize(N + 3);
    m.cell_faces.resize(2 * N + 2);
    m.cell_connects.resize(2 * N + 2);
    for (uint32_t i = 0; i < N; ++i) {
        m.cell_faces_starts[i] = 2 * i;
        m.cell_faces[2 * i] = i;
        m.cell_faces[2 * i + 1] = i + 1;
        m.cell_connects[2 * i] = i == 0 ? N : i - 1;
        m.cell_connects[2 * i + 1] = i == N - 1 ? N + 1 : i + 1;
    }
    m.cell_faces_starts[N] = 2 * N;
    m.cell_faces_starts[N + 1] = 2 * N + 1;
    m.cell_faces_starts[N + 2] = 2 * N + 2;
    m.cell_faces[2 * N] = 0;
    m.cell_faces[2 * N + 1] = N;
    m.cell_connects[2 * N] = 0;
    m.cell_connects[2 * N + 1] = N - 1;
    m.ghost_cell_ids = {N, N + 1};
    m.ghost_cell_starts = {0, 1, 2};
    m.region_names = {"x0", "x1"};
    m.one_dimensional = true;
    m.validate();
    return m;
}

// ---- .fvm binary format (src/meshio.cu:166-271 read, :1150-1216 write) ------------------

namespace {
struct File {
    FILE *f;
    explicit File(FILE *p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};
}  // namespace

void read_fvm(const char *path, int sb, HostMesh &m) {
    require(sb == 4 || sb == 8, "fvm: scalar_bytes must be 4 or 8");
    File fp(std::fopen(path, "rb"));
    require(fp.f != nullptr, std::string("cannot open fvm file ") + path, GR_ERR_IO);
    auto rd = [&](void *p, size_t bytes) {
        require(std::fread(p, 1, bytes, fp.f) == bytes, "fvm: truncated file", GR_ERR_IO);
    };
    auto u32 = [&]() { uint32_t v; rd(&v, 4); return v; };
    m = HostMesh();
    m.n_nodes = u32();
    m.nodes.resize(3ull * m.n_nodes);
    for (uint32_t i = 0; i < m.n_nodes; ++i) {
        if (sb == 8) {
            rd(&m.nodes[3ull * i], 24);
        } else {
            float x[3];
            rd(x, 12);
            for (int k = 0; k < 3; ++k) m.nodes[3ull * i + k] = x[k];
        }
    }
    m.n_faces = u32();
    uint32_t n_face_tags = u32();
    m.face_nodes_starts.resize((size_t)m.n_faces + 1);
    m.face_nodes.reserve(n_face_tags);
    for (uint32_t f = 0; f < m.n_faces; ++f) {
        uint32_t sz = u32();
        require(sz <= 64 && m.face_nodes.size() + sz <= n_face_tags, "fvm: bad face size", GR_ERR_IO);
        m.face_nodes_starts[f] = (uint32_t)m.face_nodes.size();
        for (uint32_t j = 0; j < sz; ++j) m.face_nodes.push_back(u32());
    }
    m.face_nodes_starts[m.n_faces] = (uint32_t)m.face_nodes.size();
    require(m.face_nodes.size() == n_face_tags, "fvm: face tag count mismatch", GR_ERR_IO);
    uint32_t n_cg = u32(), n_cell_tags = u32();
    m.cell_faces_starts.resize((size_t)n_cg + 1);
    m.cell_faces.reserve(n_cell_tags);
    m.cell_connects.reserve(n_cell_tags);
    for (uint32_t c = 0; c < n_cg; ++c) {
        uint32_t sz = u32();
        require(sz <= 64 && m.cell_faces.size() + sz <= n_cell_tags, "fvm: bad cell size", GR_ERR_IO);
        m.cell_faces_starts[c] = (uint32_t)m.cell_faces.size();
        for (uint32_t j = 0; j < sz; ++j) m.cell_faces.push_back(u32());
        for (uint32_t j = 0; j < sz; ++j) m.cell_connects.push_back(u32());
    }
    m.cell_faces_starts[n_cg] = (uint32_t)m.cell_faces.size();
    require(m.cell_faces.size() == n_cell_tags, "fvm: cell tag count mismatch", GR_ERR_IO);
    m.n_regions = u32();
    m.n_ghosts = u32();
    require(m.n_ghosts <= n_cg, "fvm: more ghosts than cells", GR_ERR_IO);
    m.n_cells = n_cg - m.n_ghosts;
    m.ghost_cell_starts.resize((size_t)m.n_regions + 1);
    for (uint32_t r = 0; r < m.n_regions; ++r) {
        char name[129];
        rd(name, 128);
        name[128] = 0;
        m.region_names.emplace_back(name);
        uint32_t cnt = u32();
        m.ghost_cell_starts[r] = (uint32_t)m.ghost_cell_ids.size();
        require(m.ghost_cell_ids.size() + cnt <= m.n_ghosts, "fvm: region ghost count mismatch", GR_ERR_IO);
        for (uint32_t i = 0; i < cnt; ++i) m.ghost_cell_ids.push_back(u32());
    }
    m.ghost_cell_starts[m.n_regions] = (uint32_t)m.ghost_cell_ids.size();
    require(m.ghost_cell_ids.size() == m.n_ghosts, "fvm: ghost count mismatch", GR_ERR_IO);
    m.validate();
}

void write_fvm(const HostMesh &m, const char *path, int sb) {
    require(sb == 4 || sb == 8, "fvm: scalar_bytes must be 4 or 8");
    File fp(std::fopen(path, "wb"));
    require(fp.f != nullptr, std::string("cannot open for writing ") + path, GR_ERR_IO);
    bool ok = true;
    auto wr = [&](const void *p, size_t bytes) { ok = ok && std::fwrite(p, 1, bytes, fp.f) == bytes; };
    auto u32 = [&](uint32_t v) { wr(&v, 4); };
    u32(m.n_nodes);
    for (uint32_t i = 0; i < m.n_nodes; ++i) {
        if (sb == 8) {
            wr(&m.nodes[3ull * i], 24);
        } else {
            float x[3] = {(float)m.nodes[3ull * i], (float)m.nodes[3ull * i + 1], (float)m.nodes[3ull * i + 2]};
            wr(x, 12);
        }
    }
    u32(m.n_faces);
    u32(m.face_nodes_starts[m.n_faces]);
    for (uint32_t f = 0; f < m.n_faces; ++f) {
        u32(m.face_nodes_starts[f + 1] - m.face_nodes_starts[f]);
        wr(&m.face_nodes[m.face_nodes_starts[f]], 4ull * (m.face_nodes_starts[f + 1] - m.face_nodes_starts[f]));
    }
    const uint32_t NG = m.n_cells + m.n_ghosts;
    u32(NG);
    u32(m.cell_faces_starts[NG]);
    for (uint32_t c = 0; c < NG; ++c) {
        uint32_t s = m.cell_faces_starts[c], e = m.cell_faces_starts[c + 1];
        u32(e - s);
        wr(&m.cell_faces[s], 4ull * (e - s));
        wr(&m.cell_connects[s], 4ull * (e - s));
    }
    u32(m.n_regions);
    u32(m.n_ghosts);
    for (uint32_t r = 0; r < m.n_regions; ++r) {
        char name[128];
        std::memset(name, 0, sizeof name);
        std::strncpy(name, m.region_names[r].c_str(), 127);
        wr(name, 128);
        u32(m.ghost_cell_starts[r + 1] - m.ghost_cell_starts[r]);
        for (uint32_t g = m.ghost_cell_starts[r]; g < m.ghost_cell_starts[r + 1]; ++g) u32(m.ghost_cell_ids[g]);
    }
    ok = ok && std::fflush(fp.f) == 0;
    require(ok, std::string("fvm: write failed ") + path, GR_ERR_IO);
}

// ---- renumbering ------------------------------------------------------------------------

// Reverse Cuthill-McKee order of the interior cells (new -> old).
static std::vector<uint32_t> rcm_order(const HostMesh &m) {
    const uint32_t N = m.n_cells;
    auto nbr_begin = [&](uint32_t c) { return m.cell_faces_starts[c]; };
    auto nbr_end = [&](uint32_t c) { return m.cell_faces_starts[c + 1]; };
    std::vector<uint32_t> deg(N, 0);
    for (uint32_t c = 0; c < N; ++c)
        for (uint32_t q = nbr_begin(c); q < nbr_end(c); ++q)
            if (m.cell_connects[q] < N) deg[c]++;
    std::vector<int32_t> level(N, -1);
    // BFS from s restricted to unvisited-by-ordering cells; returns the last level's cells.
    std::vector<uint8_t> placed(N, 0);
    std::vector<uint32_t> bfs_q;
    auto bfs_levels = [&](uint32_t s, int &depth) {
        bfs_q.clear();
        std::vector<uint32_t> touched;
        level[s] = 0;
        bfs_q.push_back(s);
        touched.push_back(s);
        size_t head = 0;
        depth = 0;
        while (head < bfs_q.size()) {
            uint32_t c = bfs_q[head++];
            depth = std::max(depth, (int)level[c]);
            for (uint32_t q = nbr_begin(c); q < nbr_end(c); ++q) {
                uint32_t j = m.cell_connects[q];
                if (j < N && level[j] < 0 && !placed[j]) {
                    level[j] = level[c] + 1;
                    bfs_q.push_back(j);
                    touched.push_back(j);
                }
            }
        }
        std::vector<uint32_t> last;
        for (uint32_t c : touched)
            if (level[c] == depth) last.push_back(c);
        for (uint32_t c : touched) level[c] = -1;
        return last;
    };
    std::vector<uint32_t> order;
    order.reserve(N);
    std::vector<uint32_t> nbrs;
    for (uint32_t seed = 0; seed < N; ++seed) {
        if (placed[seed]) continue;
        // pseudo-peripheral start (George-Liu): the minimum-degree cell of the last BFS level,
        // repeated while the eccentricity grows.
        uint32_t s = seed;
        int depth = 0;
        auto last = bfs_levels(s, depth);
        for (int it = 0; it < 8; ++it) {
            uint32_t best = last[0];
            for (uint32_t c : last)
                if (deg[c] < deg[best] || (deg[c] == deg[best] && c < best)) best = c;
            int d2 = 0;
            auto last2 = bfs_levels(best, d2);
            if (d2 <= depth) break;
            s = best;
            depth = d2;
            last = std::move(last2);
        }
        // Cuthill-McKee from s: neighbours by increasing degree, ties by original id
        size_t head = order.size();
        order.push_back(s);
        placed[s] = 1;
        while (head < order.size()) {
            uint32_t c = order[head++];
            nbrs.clear();
            for (uint32_t q = nbr_begin(c); q < nbr_end(c); ++q) {
                uint32_t j = m.cell_connects[q];
                if (j < N && !placed[j]) {
                    placed[j] = 1;
                    nbrs.push_back(j);
                }
            }
            std::sort(nbrs.begin(), nbrs.end(), [&](uint32_t a, uint32_t b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            order.insert(order.end(), nbrs.begin(), nbrs.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Hilbert index of a point on a 2^b grid per axis (b <= 21), by Skilling's transpose method
// ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004): axes to transposed Hilbert
// coordinates, then the bits interleaved from the most significant.
static uint64_t hilbert_key(uint32_t x0, uint32_t x1, uint32_t x2, int b) {
    uint32_t X[3] = {x0, x1, x2};
    const uint32_t M = 1u << (b - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int bit = b - 1; bit >= 0; --bit)
        for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> bit) & 1u);
    return key;
}

// Interior cells in Hilbert-curve order of their node-averaged centres (ties by id), on a
// 2^21 grid over the mesh's bounding box: cells close along the curve are close in space, so
// a contiguous run of cells (the tiles one XCD has in flight) is a compact region whose
// neighbours are few, where reverse Cuthill-McKee gives a thin slab whose neighbours span two
// bands of its bandwidth.
static std::vector<uint32_t> hilbert_order(const HostMesh &m) {
    const uint32_t N = m.n_cells;
    std::vector<double> c(3 * (size_t)N, 0.0);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = 0; i < N; ++i) {
        double s[3] = {0, 0, 0};
        uint32_t n = 0;
        for (uint32_t q = m.cell_faces_starts[i]; q < m.cell_faces_starts[i + 1]; ++q) {
            const uint32_t f = m.cell_faces[q];
            for (uint32_t k = m.face_nodes_starts[f]; k < m.face_nodes_starts[f + 1]; ++k) {
                const uint32_t v = m.face_nodes[k];
                for (int d = 0; d < 3; ++d) s[d] += m.nodes[3 * (size_t)v + d];
                ++n;
            }
        }
        for (int d = 0; d < 3; ++d) {
            c[3 * (size_t)i + d] = n ? s[d] / n : 0.0;
            lo[d] = std::min(lo[d], c[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], c[3 * (size_t)i + d]);
        }
    }
    const int bits = 21;
    const double cells = (double)((1u << bits) - 1);
    double ext = 0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    const double scale = ext > 0 ? cells / ext : 0.0;   // one scale for all axes: the curve stays isotropic
    std::vector<std::pair<uint64_t, uint32_t>> keys(N);
    for (uint32_t i = 0; i < N; ++i) {
        uint32_t g[3];
        for (int d = 0; d < 3; ++d) {
            const double v = (c[3 * (size_t)i + d] - lo[d]) * scale;
            g[d] = (uint32_t)std::min(cells, std::max(0.0, v));
        }
        keys[i] = {hilbert_key(g[0], g[1], g[2], bits), i};
    }
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> order(N);
    for (uint32_t i = 0; i < N; ++i) order[i] = keys[i].second;
    return order;
}

// Applies an interior-cell order (new -> old): ghosts region by region by their interior cell's
// new id, faces by (lower new cell, higher new cell); each cell keeps its face-slot order.
static void apply_cell_order(HostMesh &m, const std::vector<uint32_t> &order, std::vector<uint32_t> &cell_new2old,
                             std::vector<uint32_t> &face_new2old) {
    const uint32_t N = m.n_cells, G = m.n_ghosts, NG = N + G;
    auto nbr_begin = [&](uint32_t c) { return m.cell_faces_starts[c]; };
    auto nbr_end = [&](uint32_t c) { return m.cell_faces_starts[c + 1]; };
    cell_new2old.assign(NG, 0);
    std::vector<uint32_t> cell_old2new(NG);
    for (uint32_t i = 0; i < N; ++i) {
        cell_new2old[i] = order[i];
        cell_old2new[order[i]] = i;
    }
    // ghosts: region by region, ordered by their interior cell's new id
    for (uint32_t r = 0; r < m.n_regions; ++r) {
        uint32_t a = m.ghost_cell_starts[r], b = m.ghost_cell_starts[r + 1];
        std::vector<uint32_t> gs;
        for (uint32_t g = a; g < b; ++g) gs.push_back(N + g);
        std::stable_sort(gs.begin(), gs.end(), [&](uint32_t x, uint32_t y) {
            return cell_old2new[m.cell_connects[m.cell_faces_starts[x]]] <
                   cell_old2new[m.cell_connects[m.cell_faces_starts[y]]];
        });
        for (uint32_t k = 0; k < gs.size(); ++k) {
            cell_new2old[N + a + k] = gs[k];
            cell_old2new[gs[k]] = N + a + k;
        }
    }
    // faces by (lower new owner, higher new owner)
    std::vector<uint32_t> f_lo(m.n_faces, 0xFFFFFFFFu), f_hi(m.n_faces, 0);
    for (uint32_t c = 0; c < NG; ++c)
        for (uint32_t q = nbr_begin(c); q < nbr_end(c); ++q) {
            uint32_t f = m.cell_faces[q], nc = cell_old2new[c];
            f_lo[f] = std::min(f_lo[f], nc);
            f_hi[f] = std::max(f_hi[f], nc);
        }
    face_new2old.resize(m.n_faces);
    std::iota(face_new2old.begin(), face_new2old.end(), 0u);
    std::sort(face_new2old.begin(), face_new2old.end(), [&](uint32_t a, uint32_t b) {
        if (f_lo[a] != f_lo[b]) return f_lo[a] < f_lo[b];
        if (f_hi[a] != f_hi[b]) return f_hi[a] < f_hi[b];
        return a < b;
    });
    std::vector<uint32_t> face_old2new(m.n_faces);
    for (uint32_t f = 0; f < m.n_faces; ++f) face_old2new[face_new2old[f]] = f;

    HostMesh r;
    r.n_nodes = m.n_nodes;
    r.n_faces = m.n_faces;
    r.n_cells = N;
    r.n_ghosts = G;
    r.n_regions = m.n_regions;
    r.nodes = m.nodes;
    r.region_names = m.region_names;
    r.ghost_cell_starts = m.ghost_cell_starts;
    r.ghost_cell_ids = m.ghost_cell_ids;
    r.face_nodes_starts.resize(m.n_faces + 1);
    r.face_nodes_starts[0] = 0;
    for (uint32_t f = 0; f < m.n_faces; ++f) {
        uint32_t o = face_new2old[f];
        uint32_t len = m.face_nodes_starts[o + 1] - m.face_nodes_starts[o];
        r.face_nodes_starts[f + 1] = r.face_nodes_starts[f] + len;
        r.face_nodes.insert(r.face_nodes.end(), m.face_nodes.begin() + m.face_nodes_starts[o],
                            m.face_nodes.begin() + m.face_nodes_starts[o + 1]);
    }
    r.cell_faces_starts.resize(NG + 1);
    r.cell_faces_starts[0] = 0;
    r.cell_faces.reserve(m.cell_faces.size());
    r.cell_connects.reserve(m.cell_connects.size());
    for (uint32_t c = 0; c < NG; ++c) {
        uint32_t o = cell_new2old[c];
        for (uint32_t q = nbr_begin(o); q < nbr_end(o); ++q) {
            r.cell_faces.push_back(face_old2new[m.cell_faces[q]]);
            r.cell_connects.push_back(cell_old2new[m.cell_connects[q]]);
        }
        r.cell_faces_starts[c + 1] = (uint32_t)r.cell_faces.size();
    }
    m = std::move(r);
}

void renumber_rcm(HostMesh &m, std::vector<uint32_t> &cell_new2old, std::vector<uint32_t> &face_new2old) {
    apply_cell_order(m, rcm_order(m), cell_new2old, face_new2old);
}

void renumber_hilbert(HostMesh &m, std::vector<uint32_t> &cell_new2old, std::vector<uint32_t> &face_new2old) {
    apply_cell_order(m, hilbert_order(m), cell_new2old, face_new2old);
}

std::vector<uint32_t> face_colors(const HostMesh &m, uint32_t &n_colors) {
    const uint32_t NG = m.n_cells + m.n_ghosts;
    std::vector<uint32_t> owner_a(m.n_faces, 0xFFFFFFFFu), owner_b(m.n_faces, 0xFFFFFFFFu);
    for (uint32_t c = 0; c < NG; ++c)
        for (uint32_t q = m.cell_faces_starts[c]; q < m.cell_faces_starts[c + 1]; ++q) {
            uint32_t f = m.cell_faces[q];
            if (owner_a[f] == 0xFFFFFFFFu) owner_a[f] = c;
            else owner_b[f] = c;
        }
    std::vector<uint64_t> used(NG, 0);
    std::vector<uint32_t> col(m.n_faces);
    n_colors = 0;
    for (uint32_t f = 0; f < m.n_faces; ++f) {
        uint64_t mask = used[owner_a[f]] | (owner_b[f] != 0xFFFFFFFFu ? used[owner_b[f]] : 0);
        require(~mask != 0, "face colouring: more than 64 colours needed", GR_ERR_TOPOLOGY);
        uint32_t c = (uint32_t)__builtin_ctzll(~mask);
        col[f] = c;
        used[owner_a[f]] |= 1ull << c;
        if (owner_b[f] != 0xFFFFFFFFu) used[owner_b[f]] |= 1ull << c;
        n_colors = std::max(n_colors, c + 1);
    }
    return col;
}

uint32_t bandwidth(const HostMesh &m) {
    uint32_t bw = 0;
    for (uint32_t c = 0; c < m.n_cells; ++c)
        for (uint32_t q = m.cell_faces_starts[c]; q < m.cell_faces_starts[c + 1]; ++q) {
            uint32_t j = m.cell_connects[q];
            if (j < m.n_cells) bw = std::max(bw, j > c ? j - c : c - j);
        }
    return bw;
}

}  // namespace gr
