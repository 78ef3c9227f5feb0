"""One GPU's share of a mesh partitioned by contiguous ranges of (renumbered) cells.

The multi-GPU layout of the residual path: each GPU owns a contiguous range of cells in the
mesh's Hilbert (or RCM) order, plus one layer of halo values for the faces it shares with other
ranges, stored like ghosts after its own ghosts so that the iteration kernels run unchanged.
`shard_elements` cuts that share out of an element list: the range's volume elements, its
boundary faces under their own markers, and the cut faces as one more marker whose ghosts carry
no boundary condition (the kernels read their values from the state array, where a halo exchange
would write them).  Host-side numpy; used to time the kernels at one GPU's real share.
"""
from __future__ import annotations

import numpy as np

from . import GR_QUAD, GR_TRI, Elements, Mesh


def _face_lists(con: dict, faces: np.ndarray):
    """(kinds, starts, nodes) of mesh faces `faces` as marker elements."""
    fns, fn = con["face_nodes_starts"].astype(np.int64), con["face_nodes"]
    a, b = fns[faces], fns[faces + 1]
    lens = b - a
    if lens.size and not np.all((lens == 3) | (lens == 4)):
        raise ValueError("faces must be triangles or quads")
    kinds = np.where(lens == 3, GR_TRI, GR_QUAD).astype(np.uint32)
    starts = np.zeros(len(faces) + 1, np.uint32)
    np.cumsum(lens, out=starts[1:])
    idx = np.repeat(a - starts[:-1], lens) + np.arange(int(starts[-1]))
    return kinds, starts, fn[idx].astype(np.uint32)


def shard_elements(e: Elements, m: Mesh, c2o: np.ndarray, lo: int, hi: int, halo: str = "halo") -> Elements:
    """Elements of cells lo..hi-1 of `m`, a mesh built from `e` and renumbered with cell map c2o
    (new index -> index in `e`, as `Mesh.renumber_hilbert` / `renumber_rcm` return it).  Markers:
    the boundary faces of those cells under the name of their region, then the faces to cells
    outside the range under `halo` (omitted when there are none).  Node coordinates are kept
    whole (unreferenced nodes are allowed)."""
    s = m.sizes()
    n = s.n_cells
    if not 0 <= lo < hi <= n:
        raise ValueError(f"cell range [{lo}, {hi}) outside 0..{n}")
    nodes, kinds, starts, enodes, _ = e.arrays()
    con = m.connectivity()
    # volume elements of the range, in the range's order
    old = c2o[lo:hi].astype(np.int64)
    st = starts.astype(np.int64)
    lens = st[old + 1] - st[old]
    k2 = kinds[old].astype(np.uint32)
    s2 = np.zeros(len(old) + 1, np.uint32)
    np.cumsum(lens, out=s2[1:])
    n2 = enodes[np.repeat(st[old] - s2[:-1].astype(np.int64), lens) + np.arange(int(s2[-1]))]
    # the range's face slots: boundary faces by region, cut faces
    cfs = con["cell_faces_starts"].astype(np.int64)
    q = np.arange(cfs[lo], cfs[hi])
    j = con["cell_connects"][q].astype(np.int64)
    f = con["cell_faces"][q].astype(np.int64)
    gcs = con["ghost_cell_starts"].astype(np.int64)
    markers = []
    ghost = j >= n
    region = np.searchsorted(gcs, j[ghost] - n, side="right") - 1
    for r, name in enumerate(m.region_names()):
        faces = f[ghost][region == r]
        if faces.size:
            markers.append((name, *_face_lists(con, faces)))
    cut = (~ghost) & ((j < lo) | (j >= hi))
    if cut.any():
        markers.append((halo, *_face_lists(con, f[cut])))
    return Elements.create(nodes, k2, s2, n2, markers)
