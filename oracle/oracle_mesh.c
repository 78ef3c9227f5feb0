/*
 * oracle_mesh.c -- connectivity builder and mesh metrics of the FP64 CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The connectivity builder is a literal chained hash table, built the way the reference's
 * cuMeshFromSu2 builds it (src/meshio.cu:713-764, :902-1138): buckets are visited in index
 * order and each chain in creation order, which fixes the face numbering.  The product
 * builder (graphite_amd/csrc/connectivity.cpp) reaches the same numbering with a sort, so
 * the two are independent implementations of one specification.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Local face tables of the su2 element kinds, src/meshio.cu:297-475.  Each face lists up to
 * four local node indices, -1 padded. */
typedef struct { int n_nodes; int n_faces; int f[6][4]; } elem_kind;

static int kind_info(uint32_t kind, elem_kind *k)
{
    static const elem_kind TRI = {3, 1, {{0, 1, 2, -1}}};
    static const elem_kind QUAD = {4, 1, {{0, 1, 2, 3}}};
    static const elem_kind TET = {4, 4, {{0, 2, 1, -1}, {0, 1, 3, -1}, {1, 2, 3, -1}, {0, 3, 2, -1}}};
    static const elem_kind HEX = {8, 6, {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7},
                                         {3, 2, 1, 0}, {4, 5, 6, 7}}};
    static const elem_kind WEDGE = {6, 5, {{0, 1, 2, -1}, {3, 5, 4, -1}, {0, 2, 5, 3}, {1, 4, 5, 2},
                                           {1, 0, 3, 4}}};
    static const elem_kind PYR = {5, 5, {{0, 1, 2, 3}, {0, 4, 3, -1}, {0, 1, 4, -1}, {1, 2, 4, -1},
                                         {2, 3, 4, -1}}};
    switch (kind) {
    case 5: *k = TRI; return 0;
    case 9: *k = QUAD; return 0;
    case 10: *k = TET; return 0;
    case 12: *k = HEX; return 0;
    case 13: *k = WEDGE; return 0;
    case 14: *k = PYR; return 0;
    default: return 1;
    }
}

static void sort_small(uint32_t *a, uint32_t n)
{
    for (uint32_t i = 1; i < n; ++i) {
        uint32_t v = a[i];
        uint32_t j = i;
        while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; --j; }
        a[j] = v;
    }
}

uint32_t orc_face_hash(const uint32_t *key, uint32_t n, uint32_t table_size)
{
    uint32_t s[8];
    memcpy(s, key, n * sizeof(uint32_t));
    sort_small(s, n);
    uint32_t v = 1u;
    for (uint32_t i = 0; i < n; ++i) v *= 998244353u + 2u * s[i];
    return (v / 2u) % table_size;
}

#define ORC_MAXVAL 8
typedef struct {
    uint32_t key[4];
    uint32_t sorted[4];
    uint32_t ksize;
    uint32_t vals[ORC_MAXVAL];
    uint32_t nvals;       /* may exceed ORC_MAXVAL on broken meshes; only the count is used */
    int64_t next;
} chain_item;

typedef struct {
    uint32_t size;
    int64_t *head;
    int64_t *tail;
    chain_item *items;
    size_t n_items, cap;
} chain_table;

static int same_face(const chain_item *it, const uint32_t *sorted, uint32_t n)
{
    if (it->ksize != n) return 0;
    for (uint32_t i = 0; i < n; ++i)
        if (it->sorted[i] != sorted[i]) return 0;
    return 1;
}

static void add_val(chain_item *it, uint32_t v)
{
    if (it->nvals < ORC_MAXVAL) it->vals[it->nvals] = v;
    it->nvals++;
}

static int64_t table_find(const chain_table *t, const uint32_t *key, uint32_t n)
{
    uint32_t s[4];
    memcpy(s, key, n * sizeof(uint32_t));
    sort_small(s, n);
    uint32_t b = orc_face_hash(key, n, t->size);
    for (int64_t i = t->head[b]; i >= 0; i = t->items[i].next)
        if (same_face(&t->items[i], s, n)) return i;
    return -1;
}

/* FaceHashMap_add semantics: value goes to the matching item of the chain, else a new item is
 * appended at the tail of the chain. */
static void table_add(chain_table *t, const uint32_t *key, uint32_t n, uint32_t value)
{
    int64_t found = table_find(t, key, n);
    if (found >= 0) { add_val(&t->items[found], value); return; }
    uint32_t b = orc_face_hash(key, n, t->size);
    chain_item *it = &t->items[t->n_items];
    memset(it, 0, sizeof(*it));
    memcpy(it->key, key, n * sizeof(uint32_t));
    memcpy(it->sorted, key, n * sizeof(uint32_t));
    sort_small(it->sorted, n);
    it->ksize = n;
    it->next = -1;
    add_val(it, value);
    int64_t idx = (int64_t)t->n_items++;
    if (t->head[b] < 0) t->head[b] = idx;
    else t->items[t->tail[b]].next = idx;
    t->tail[b] = idx;
}

/* Nodes of local face fi of an element, in the element's local order. */
static uint32_t face_of(const elem_kind *k, const uint32_t *en, int fi, uint32_t *out)
{
    uint32_t n = 0;
    for (int j = 0; j < 4; ++j)
        if (k->f[fi][j] >= 0) out[n++] = en[k->f[fi][j]];
    return n;
}

int orc_build_mesh(uint32_t n_nodes, const double *nodes,
                   uint32_t n_elems, const uint32_t *elem_types,
                   const uint32_t *elem_starts, const uint32_t *elem_nodes,
                   uint32_t n_markers, const uint32_t *marker_counts,
                   const uint32_t *marker_types, const uint32_t *marker_starts,
                   const uint32_t *marker_nodes, orc_mesh **out)
{
    *out = NULL;
    int rc = 1;
    uint32_t n_ghosts = 0;
    for (uint32_t m = 0; m < n_markers; ++m) n_ghosts += marker_counts[m];
    const uint32_t n_total = n_elems + n_ghosts;

    /* element faces, src/meshio.cu:713-764 */
    size_t cap = 0;
    for (uint32_t e = 0; e < n_elems; ++e) {
        elem_kind k;
        if (kind_info(elem_types[e], &k)) return 1;
        cap += (size_t)k.n_faces;
    }
    chain_table t;
    t.size = n_elems + 1;
    t.head = (int64_t *)malloc(sizeof(int64_t) * t.size);
    t.tail = (int64_t *)malloc(sizeof(int64_t) * t.size);
    t.items = (chain_item *)malloc(sizeof(chain_item) * (cap ? cap : 1));
    t.n_items = 0;
    t.cap = cap;
    for (uint32_t b = 0; b < t.size; ++b) t.head[b] = t.tail[b] = -1;

    uint32_t *ghost_region = (uint32_t *)malloc(sizeof(uint32_t) * (n_ghosts ? n_ghosts : 1));
    uint32_t *ghost_local = (uint32_t *)malloc(sizeof(uint32_t) * (n_ghosts ? n_ghosts : 1));
    uint32_t *marker_node_off = (uint32_t *)malloc(sizeof(uint32_t) * (n_markers + 1));
    uint32_t *cell_sizes = (uint32_t *)calloc(n_total ? n_total : 1, sizeof(uint32_t));
    uint32_t *fid_of_item = NULL;
    orc_mesh *m = NULL;

    for (uint32_t e = 0; e < n_elems; ++e) {
        elem_kind k;
        kind_info(elem_types[e], &k);
        const uint32_t *en = elem_nodes + elem_starts[e];
        for (int fi = 0; fi < k.n_faces; ++fi) {
            uint32_t fn[4];
            uint32_t n = face_of(&k, en, fi, fn);
            table_add(&t, fn, n, e);
        }
    }

    /* ghost ids n_elems + running marker offset, attached to their face (:978-1007) */
    {
        uint32_t g = 0, noff = 0, soff = 0;
        for (uint32_t mk = 0; mk < n_markers; ++mk) {
            marker_node_off[mk] = noff;
            const uint32_t *ms = marker_starts + soff;
            for (uint32_t i = 0; i < marker_counts[mk]; ++i, ++g) {
                const uint32_t *mn = marker_nodes + noff + ms[i];
                uint32_t n = ms[i + 1] - ms[i];
                if (n > 4) goto fail;
                int64_t it = table_find(&t, mn, n);
                if (it < 0) goto fail;
                add_val(&t.items[it], n_elems + g);
                ghost_region[g] = mk;
                ghost_local[g] = i;
            }
            noff += ms[marker_counts[mk]];
            soff += marker_counts[mk] + 1;
        }
        marker_node_off[n_markers] = noff;
    }

    m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->n_nodes = n_nodes;
    m->n_cells = n_elems;
    m->n_ghosts = n_ghosts;
    m->n_regions = n_markers;
    m->n_faces = (uint32_t)t.n_items;

    /* face arrays in bucket order, then chain order (:920-956) */
    fid_of_item = (uint32_t *)malloc(sizeof(uint32_t) * (t.n_items ? t.n_items : 1));
    {
        size_t total_fn = 0;
        for (size_t i = 0; i < t.n_items; ++i) total_fn += t.items[i].ksize;
        m->face_nodes_starts = (uint32_t *)malloc(sizeof(uint32_t) * (m->n_faces + 1));
        m->face_nodes = (uint32_t *)malloc(sizeof(uint32_t) * (total_fn ? total_fn : 1));
        uint32_t fid = 0, off = 0;
        for (uint32_t b = 0; b < t.size; ++b) {
            for (int64_t i = t.head[b]; i >= 0; i = t.items[i].next) {
                chain_item *it = &t.items[i];
                m->face_nodes_starts[fid] = off;
                for (uint32_t j = 0; j < it->ksize; ++j) m->face_nodes[off++] = it->key[j];
                /* every owner gets one more face (:1015-1031) */
                uint32_t nv = it->nvals < ORC_MAXVAL ? it->nvals : ORC_MAXVAL;
                for (uint32_t j = 0; j < nv; ++j) cell_sizes[it->vals[j]]++;
                fid_of_item[i] = fid;
                add_val(it, fid);
                fid++;
            }
        }
        m->face_nodes_starts[m->n_faces] = off;
    }

    /* cell CSR in element-local face order (:1032-1136) */
    m->cell_faces_starts = (uint32_t *)malloc(sizeof(uint32_t) * (n_total + 1));
    {
        uint32_t acc = 0;
        for (uint32_t c = 0; c < n_total; ++c) { m->cell_faces_starts[c] = acc; acc += cell_sizes[c]; }
        m->cell_faces_starts[n_total] = acc;
        m->cell_faces = (uint32_t *)malloc(sizeof(uint32_t) * (acc ? acc : 1));
        m->cell_connects = (uint32_t *)malloc(sizeof(uint32_t) * (acc ? acc : 1));
    }
    {
        uint32_t off = 0;
        for (uint32_t c = 0; c < n_total; ++c) {
            elem_kind k;
            const uint32_t *en;
            if (c < n_elems) {
                kind_info(elem_types[c], &k);
                en = elem_nodes + elem_starts[c];
            } else {
                uint32_t g = c - n_elems, mk = ghost_region[g], i = ghost_local[g];
                uint32_t soff = 0;
                for (uint32_t q = 0; q < mk; ++q) soff += marker_counts[q] + 1;
                const uint32_t *ms = marker_starts + soff;
                if (kind_info(marker_types[g], &k)) goto fail;
                en = marker_nodes + marker_node_off[mk] + ms[i];
            }
            for (int fi = 0; fi < k.n_faces; ++fi) {
                uint32_t fn[4];
                uint32_t n = face_of(&k, en, fi, fn);
                int64_t it = table_find(&t, fn, n);
                if (it < 0) goto fail;
                chain_item *item = &t.items[it];
                if (item->nvals != 3) goto fail;   /* two owners plus the face id */
                if (off >= m->cell_faces_starts[n_total]) goto fail;
                m->cell_faces[off] = item->vals[2];
                m->cell_connects[off] = (item->vals[0] == c) ? item->vals[1] : item->vals[0];
                off++;
            }
        }
    }

    /* regions: ghosts are n_elems.. in marker order (:1194-1213) */
    m->ghost_cell_ids = (uint32_t *)malloc(sizeof(uint32_t) * (n_ghosts ? n_ghosts : 1));
    m->ghost_cell_starts = (uint32_t *)malloc(sizeof(uint32_t) * (n_markers + 1));
    {
        uint32_t g = 0;
        for (uint32_t mk = 0; mk < n_markers; ++mk) {
            m->ghost_cell_starts[mk] = g;
            for (uint32_t i = 0; i < marker_counts[mk]; ++i, ++g) m->ghost_cell_ids[g] = n_elems + g;
        }
        m->ghost_cell_starts[n_markers] = g;
    }

    m->nodes = (double *)malloc(sizeof(double) * 3 * (n_nodes ? n_nodes : 1));
    memcpy(m->nodes, nodes, sizeof(double) * 3 * n_nodes);
    size_t ncg = (size_t)n_total;
    m->face_centers = (double *)calloc(3 * (size_t)m->n_faces + 1, sizeof(double));
    m->face_normals = (double *)calloc(3 * (size_t)m->n_faces + 1, sizeof(double));
    m->face_areas = (double *)calloc((size_t)m->n_faces + 1, sizeof(double));
    m->cell_centers = (double *)calloc(3 * ncg + 1, sizeof(double));
    m->cell_volumes = (double *)calloc(ncg + 1, sizeof(double));
    m->cell_matrices = (double *)calloc(9 * (size_t)n_elems + 1, sizeof(double));
    *out = m;
    m = NULL;
    rc = 0;

fail:
    if (m) orc_mesh_free(m);
    free(fid_of_item);
    free(cell_sizes);
    free(marker_node_off);
    free(ghost_local);
    free(ghost_region);
    free(t.items);
    free(t.tail);
    free(t.head);
    return rc;
}

void orc_mesh_free(orc_mesh *m)
{
    if (!m) return;
    free(m->nodes);
    free(m->face_nodes_starts);
    free(m->face_nodes);
    free(m->cell_faces_starts);
    free(m->cell_faces);
    free(m->cell_connects);
    free(m->ghost_cell_ids);
    free(m->ghost_cell_starts);
    free(m->face_centers);
    free(m->face_normals);
    free(m->face_areas);
    free(m->cell_centers);
    free(m->cell_volumes);
    free(m->cell_matrices);
    free(m);
}

/* ---- metrics --------------------------------------------------------------------------- */

static inline double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* sign of the double build, include/datatypes.cuh:209-211: +1, 0 or -1 */
static inline double sgn(double x) { return (double)(x > 0) - (double)(x < 0); }

/* One fan triangle (centre, a, b) of a face, src/mesh.cu:67-93. */
static void fan_tri(const double *c, const double *a, const double *b, double *area,
                    double *normal, double *ncent)
{
    double d0[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double d1[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    double cr[3] = {d0[1] * d1[2] - d0[2] * d1[1], d0[2] * d1[0] - d0[0] * d1[2],
                    d0[0] * d1[1] - d0[1] * d1[0]};
    double nc = norm3(cr);
    *area += nc * 0.5;
    for (int k = 0; k < 3; ++k) {
        double tc = ((a[k] + b[k]) + c[k]) / 3.0;
        normal[k] += cr[k];
        ncent[k] += (tc * nc) * 0.5;
    }
}

void orc_mesh_metrics(orc_mesh *m)
{
    const uint32_t N = m->n_cells, G = m->n_ghosts;

    /* cuMeshComputeFaces, src/mesh.cu:44-97 */
    for (uint32_t f = 0; f < m->n_faces; ++f) {
        uint32_t s = m->face_nodes_starts[f], e = m->face_nodes_starts[f + 1];
        double c[3] = {0, 0, 0};
        for (uint32_t i = s; i < e; ++i)
            for (int k = 0; k < 3; ++k) c[k] += m->nodes[3 * m->face_nodes[i] + k];
        double fsz = (double)(float)(e - s);
        for (int k = 0; k < 3; ++k) c[k] /= fsz;
        double area = 0, nrm[3] = {0, 0, 0}, nc[3] = {0, 0, 0};
        for (uint32_t i = s; i + 1 < e; ++i)
            fan_tri(c, &m->nodes[3 * m->face_nodes[i]], &m->nodes[3 * m->face_nodes[i + 1]], &area, nrm, nc);
        fan_tri(c, &m->nodes[3 * m->face_nodes[e - 1]], &m->nodes[3 * m->face_nodes[s]], &area, nrm, nc);
        double nn = norm3(nrm);
        for (int k = 0; k < 3; ++k) {
            m->face_centers[3 * f + k] = nc[k] / area;
            m->face_normals[3 * f + k] = nrm[k] / nn;
        }
        m->face_areas[f] = area;
    }

    /* cuMeshComputeCells, src/mesh.cu:102-143 */
    for (uint32_t i = 0; i < N; ++i) {
        uint32_t s = m->cell_faces_starts[i], e = m->cell_faces_starts[i + 1];
        double c[3] = {0, 0, 0}, w = 0;
        for (uint32_t q = s; q < e; ++q) {
            uint32_t f = m->cell_faces[q];
            for (int k = 0; k < 3; ++k) c[k] += m->face_centers[3 * f + k] * m->face_areas[f];
            w += m->face_areas[f];
        }
        for (int k = 0; k < 3; ++k) c[k] /= w;
        double nc[3] = {0, 0, 0};
        w = 0;
        for (uint32_t q = s; q < e; ++q) {
            uint32_t f = m->cell_faces[q];
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = m->face_centers[3 * f + k] - c[k];
            double wf = norm3(d) * m->face_areas[f];
            for (int k = 0; k < 3; ++k) nc[k] += (c[k] + d[k] * 0.75) * wf;
            w += wf;
        }
        for (int k = 0; k < 3; ++k) nc[k] /= w;
        for (int k = 0; k < 3; ++k) m->cell_centers[3 * i + k] = nc[k];
        double vol = 0;
        for (uint32_t q = s; q < e; ++q) {
            uint32_t f = m->cell_faces[q];
            double d[3], n[3];
            for (int k = 0; k < 3; ++k) d[k] = m->face_centers[3 * f + k] - nc[k];
            double sg = sgn(dot3(&m->face_normals[3 * f], d));
            for (int k = 0; k < 3; ++k) n[k] = m->face_normals[3 * f + k] * sg;
            vol += dot3(d, n) * m->face_areas[f];
        }
        m->cell_volumes[i] = vol / 3.0;
    }

    /* cuMeshComputeGhostCells, src/mesh.cu:148-173 */
    for (uint32_t g = 0; g < G; ++g) {
        uint32_t i = m->ghost_cell_ids[g];
        uint32_t s = m->cell_faces_starts[i];
        uint32_t o = m->cell_connects[s], f = m->cell_faces[s];
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * o + k];
        double dx = norm3(d);
        double sg = sgn(dot3(&m->face_normals[3 * f], d));
        for (int k = 0; k < 3; ++k)
            m->cell_centers[3 * i + k] = m->face_centers[3 * f + k] + (m->face_normals[3 * f + k] * sg) * dx;
        m->cell_volumes[i] = (m->face_areas[f] * 2) * dx;
    }

    /* cuMeshComputeCellMatrices, src/mesh.cu:178-219.  A is accumulated as the tensor of
     * columns (u,v,w); entry (r,c) = d_r*d_c*w*w.  The stored matrix is adj(A)/det(A). */
    for (uint32_t i = 0; i < N; ++i) {
        double A[9] = {0};   /* A[3*col+row] */
        uint32_t s = m->cell_faces_starts[i], e = m->cell_faces_starts[i + 1];
        for (uint32_t q = s; q < e; ++q) {
            uint32_t j = m->cell_connects[q];
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = m->cell_centers[3 * j + k] - m->cell_centers[3 * i + k];
            double w = 1.0 / norm3(d);
            for (int col = 0; col < 3; ++col)
                for (int row = 0; row < 3; ++row) A[3 * col + row] += ((d[row] * d[col]) * w) * w;
        }
#define AT(col, row) A[3 * (col) + (row)]
        double M[9];
        M[0] = AT(1, 1) * AT(2, 2) - AT(1, 2) * AT(2, 1);
        M[1] = AT(2, 1) * AT(0, 2) - AT(2, 2) * AT(0, 1);
        M[2] = AT(0, 1) * AT(1, 2) - AT(0, 2) * AT(1, 1);
        M[3] = AT(2, 0) * AT(1, 2) - AT(2, 2) * AT(1, 0);
        M[4] = AT(0, 0) * AT(2, 2) - AT(0, 2) * AT(2, 0);
        M[5] = AT(1, 0) * AT(0, 2) - AT(1, 2) * AT(0, 0);
        M[6] = AT(1, 0) * AT(2, 1) - AT(1, 1) * AT(2, 0);
        M[7] = AT(2, 0) * AT(0, 1) - AT(2, 1) * AT(0, 0);
        M[8] = AT(0, 0) * AT(1, 1) - AT(0, 1) * AT(1, 0);
        double det = (AT(0, 0) * M[0] + AT(1, 0) * M[1]) + AT(2, 0) * M[2];
#undef AT
        for (int k = 0; k < 9; ++k) m->cell_matrices[9 * i + k] = M[k] / det;
    }
}

/* Mesh validity, the rule the product applies after its metric kernels (the reference runs such
 * meshes silently).  counts[0]: interior cells with a volume that is not positive and finite;
 * [1]: face slots of interior cells whose face does not separate the two cell centres, i.e.
 * n.(x_f - x_i) and n.(x_f - x_j) are not of strictly opposite signs (a face with n.d = 0, where
 * the binary64 sign() of include/datatypes.cuh:209-211 returns 0 and outer_normal drops the face,
 * or a tangled cell); [2]: interior cells with a non-finite least-squares matrix; [3]: faces
 * whose area is not positive and finite or whose normal is not finite. */
void orc_mesh_check(const orc_mesh *m, uint32_t counts[4])
{
    const uint32_t N = m->n_cells;
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    for (uint32_t i = 0; i < N; ++i) {
        const double v = m->cell_volumes[i];
        if (!(v > 0.0) || !isfinite(v)) counts[0]++;
        const double *xi = &m->cell_centers[3 * i];
        for (uint32_t q = m->cell_faces_starts[i]; q < m->cell_faces_starts[i + 1]; ++q) {
            const uint32_t f = m->cell_faces[q], j = m->cell_connects[q];
            const double *fc = &m->face_centers[3 * f], *n = &m->face_normals[3 * f];
            const double *xj = &m->cell_centers[3 * j];
            const double si = (n[0] * (fc[0] - xi[0]) + n[1] * (fc[1] - xi[1])) + n[2] * (fc[2] - xi[2]);
            const double sj = (n[0] * (fc[0] - xj[0]) + n[1] * (fc[1] - xj[1])) + n[2] * (fc[2] - xj[2]);
            if (!((si > 0.0 && sj < 0.0) || (si < 0.0 && sj > 0.0))) counts[1]++;
        }
        int fin = 1;
        for (int k = 0; k < 9; ++k) fin = fin && isfinite(m->cell_matrices[9 * (size_t)i + k]);
        if (!fin) counts[2]++;
    }
    for (uint32_t f = 0; f < m->n_faces; ++f) {
        const double a = m->face_areas[f], *n = &m->face_normals[3 * f];
        if (!(a > 0.0) || !isfinite(a) || !isfinite(n[0]) || !isfinite(n[1]) || !isfinite(n[2])) counts[3]++;
    }
}
