"""Writes tests/golden/connectivity_*.json: tiny element lists with the CSR arrays that
cuMeshFromSu2 (Alexsimulation/graphite src/meshio.cu:713-1138) produces for them.

The expected arrays come from a pure-Python walk-through of the reference's procedure,
written independently of both the C oracle and the product's sort-based builder:
  * a hash table of n_elems+1 buckets whose chains are Python lists appended at the tail
    (FaceHashMap_add, :606-664);
  * bucket = ((prod over the sorted key of (998244353 + 2*k)) mod 2**32) // 2 % size
    (facehashfunc, :551-563);
  * faces numbered by walking buckets in order and each chain front to back (:938-956);
  * ghost ids n_elems + running marker index, attached by key lookup (:978-1007);
  * each cell lists its faces in its element's local face order, with the other owner as
    neighbour (:1044-1136).
The reference ships no meshes, tests or outputs, so no reference-produced vector exists;
these fixtures pin the ordering rule on cases small enough to check by hand.

Run from the repository root: python tests/golden/make_connectivity_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# local faces per su2 kind (src/meshio.cu:315-475)
FACES = {
    5: [[0, 1, 2]],
    9: [[0, 1, 2, 3]],
    10: [[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]],
    12: [[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [3, 2, 1, 0], [4, 5, 6, 7]],
    13: [[0, 1, 2], [3, 5, 4], [0, 2, 5, 3], [1, 4, 5, 2], [1, 0, 3, 4]],
    14: [[0, 1, 2, 3], [0, 4, 3], [0, 1, 4], [1, 2, 4], [2, 3, 4]],
}
NNODES = {5: 3, 9: 4, 10: 4, 12: 8, 13: 6, 14: 5}


def bucket(key, size):
    v = 1
    for k in sorted(key):
        v = (v * (998244353 + 2 * k)) % (1 << 32)
    return (v // 2) % size


def reference_connectivity(kinds, elems, markers):
    """kinds: list of su2 kinds; elems: list of node lists; markers: [(name, [node lists])]."""
    n_elems = len(elems)
    size = n_elems + 1
    table = [[] for _ in range(size)]   # chains of items {key, vals}
    for e, (kind, en) in enumerate(zip(kinds, elems)):
        for lf in FACES[kind]:
            key = [en[j] for j in lf]
            chain = table[bucket(key, size)]
            for item in chain:
                if sorted(item["key"]) == sorted(key):
                    item["vals"].append(e)
                    break
            else:
                chain.append({"key": key, "vals": [e]})
    gid = n_elems
    for _name, faces in markers:
        for key in faces:
            chain = table[bucket(key, size)]
            item = next(it for it in chain if sorted(it["key"]) == sorted(key))
            item["vals"].append(gid)
            gid += 1
    face_nodes, face_starts = [], [0]
    fid = 0
    for chain in table:
        for item in chain:
            face_nodes += item["key"]
            face_starts.append(len(face_nodes))
            item["vals"].append(fid)
            fid += 1
    n_ghosts = gid - n_elems
    cell_faces, cell_connects, cell_starts = [], [], [0]
    all_cells = [(k, en) for k, en in zip(kinds, elems)]
    for _name, faces in markers:
        all_cells += [(5 if len(f) == 3 else 9, f) for f in faces]
    for c, (kind, en) in enumerate(all_cells):
        for lf in FACES[kind]:
            key = [en[j] for j in lf]
            chain = table[bucket(key, size)]
            item = next(it for it in chain if sorted(it["key"]) == sorted(key))
            assert len(item["vals"]) == 3, "face not shared by exactly two cells"
            cell_faces.append(item["vals"][2])
            cell_connects.append(item["vals"][1] if item["vals"][0] == c else item["vals"][0])
        cell_starts.append(len(cell_faces))
    ghost_starts = [0]
    for _name, faces in markers:
        ghost_starts.append(ghost_starts[-1] + len(faces))
    return dict(face_nodes_starts=face_starts, face_nodes=face_nodes, cell_faces_starts=cell_starts,
                cell_faces=cell_faces, cell_connects=cell_connects,
                ghost_cell_ids=list(range(n_elems, n_elems + n_ghosts)), ghost_cell_starts=ghost_starts)


def boundary_faces(kinds, elems):
    """Faces met once, in element order (the marker lists of the fixtures)."""
    count = {}
    for kind, en in zip(kinds, elems):
        for lf in FACES[kind]:
            k = tuple(sorted(en[j] for j in lf))
            count[k] = count.get(k, 0) + 1
    out = []
    for kind, en in zip(kinds, elems):
        for lf in FACES[kind]:
            key = [en[j] for j in lf]
            if count[tuple(sorted(key))] == 1:
                out.append(key)
    return out


def hex_box(nx, ny, nz):
    nodes = [[i, j, k] for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)]
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    elems = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                elems.append([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                              nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)])
    return nodes, [12] * len(elems), elems


def cube_6tets():
    # Kuhn split of the unit cube along 0-(1,1,1); corner index = x + 2y + 4z
    nodes = [[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    elems = []
    for p in perms:
        o1 = 1 << p[0]
        o2 = o1 | (1 << p[1])
        elems.append([0, o1, o2, 7])
    return nodes, [10] * 6, elems


def mixed():
    # unit cube hex, a pyramid on its top face, a wedge on its +x face, a tet on the wedge
    nodes = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
             [0.5, 0.5, 1.6], [1.7, 0, 0.5], [1.7, 1, 0.5], [2.3, 0.5, 0.4]]
    kinds = [12, 14, 13, 10]
    elems = [[0, 1, 2, 3, 4, 5, 6, 7],
             [4, 5, 6, 7, 8],
             [1, 5, 9, 2, 6, 10],
             [9, 10, 6, 11]]
    return nodes, kinds, elems


def case(name, nodes, kinds, elems, split_markers=None):
    faces = boundary_faces(kinds, elems)
    if split_markers is None:
        markers = [("wall", faces)]
    else:
        markers = split_markers(nodes, faces)
    exp = reference_connectivity(kinds, elems, markers)
    starts = [0]
    for en in elems:
        starts.append(starts[-1] + len(en))
    assert all(len(en) == NNODES[k] for k, en in zip(kinds, elems))
    return dict(name=name, nodes=nodes, kinds=kinds, elem_starts=starts,
                elem_nodes=[v for en in elems for v in en],
                markers=[dict(name=n, kinds=[5 if len(f) == 3 else 9 for f in fs],
                              starts=[sum(len(x) for x in fs[:i]) for i in range(len(fs) + 1)],
                              nodes=[v for f in fs for v in f]) for n, fs in markers],
                expected=exp)


def by_side(nodes, faces):
    # markers xmin/xmax/rest to exercise several regions
    xs = [("xmin", []), ("xmax", []), ("rest", [])]
    xmax = max(p[0] for p in nodes)
    for f in faces:
        if all(nodes[v][0] == 0 for v in f):
            xs[0][1].append(f)
        elif all(nodes[v][0] == xmax for v in f):
            xs[1][1].append(f)
        else:
            xs[2][1].append(f)
    return xs


def main():
    cases = [
        case("one_tet", [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], [10], [[0, 1, 2, 3]]),
        case("one_hex", *hex_box(1, 1, 1)),
        case("hex_2x2x2", *hex_box(2, 2, 2), split_markers=by_side),
        case("cube_6tets", *cube_6tets(), split_markers=by_side),
        case("mixed_4", *mixed()),
    ]
    for c in cases:
        path = os.path.join(HERE, f"connectivity_{c['name']}.json")
        with open(path, "w") as f:
            json.dump(c, f)
        print("wrote", path)


if __name__ == "__main__":
    main()
