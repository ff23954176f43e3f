"""Synthetic unstructured meshes in OpenFOAM polyMesh order, plus polyMesh file I/O.

The reference reads a mesh as ``points / faces / owner / neighbour / boundary``
(src/mesh_geometry_and_topology.f90:310-577).  Its conventions, which every mesh
produced here honours:

* internal faces come first, sorted by (owner, neighbour) with owner < neighbour
  ("upper-triangular order"); the face normal (right-hand rule over the node
  list) points from owner to neighbour;
* boundary faces follow, grouped into patches; their normals point outwards;
* the simplified ``boundary`` file is one ``#`` comment line followed by
  ``bctype nFaces startFace`` lines with a 0-based ``startFace``
  (src/mesh_geometry_and_topology.f90:409-470);
* the reference only ever uses the *first* ``wall`` patch's ``startFace`` and
  the total wall count (src/fvm_laplacian.f90:140-143), so wall patches must be
  contiguous -- :class:`PolyMesh` checks that.

All index arrays in :class:`PolyMesh` are 0-based int32; the C ABI takes the
reference's 1-based convention and the conversion is done at that boundary.

Nothing here computes geometry: face areas, volumes, centroids and the
interpolation factor are produced on the GPU by ``fc_mesh_geometry``.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Boundary types the reference recognises (src/mesh_geometry_and_topology.f90:417-468).
BC_TYPES = ("inlet", "outlet", "symmetry", "wall", "wallIsoth", "wallAdiab",
            "wallQFlux", "prOutlet", "domain", "cyclic")
WALL_TYPES = ("wall", "wallIsoth", "wallAdiab", "wallQFlux")


@dataclasses.dataclass
class PolyMesh:
    points: np.ndarray          # (numNodes, 3) float64
    face_ptr: np.ndarray        # (numFaces+1,) int32, CSR pointer into face_nodes
    face_nodes: np.ndarray      # (face_ptr[-1],) int32, 0-based node ids
    owner: np.ndarray           # (numFaces,) int32, 0-based
    neighbour: np.ndarray       # (numInnerFaces,) int32, 0-based
    n_cells: int
    patches: List[Tuple[str, int, int]]   # (bctype, nFaces, startFace 0-based)

    # ------------------------------------------------------------------ sizes
    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_inner(self) -> int:
        return int(self.neighbour.shape[0])

    @property
    def n_boundary(self) -> int:
        return self.n_faces - self.n_inner

    @property
    def nnz(self) -> int:
        return 2 * self.n_inner + self.n_cells

    def wall_range(self) -> Tuple[int, int]:
        """(first wall face 0-based, number of wall faces) as the reference derives them."""
        start, count = None, 0
        for bctype, nf, sf in self.patches:
            if bctype in WALL_TYPES:
                if count == 0:
                    start = sf
                count += nf
        return (0 if start is None else start), count

    def validate(self) -> None:
        nf, nif, n = self.n_faces, self.n_inner, self.n_cells
        assert self.face_ptr.shape == (nf + 1,) and self.face_ptr[0] == 0
        assert np.all(np.diff(self.face_ptr) >= 3)
        assert self.owner.min() >= 0 and self.owner.max() < n
        if nif:
            o = self.owner[:nif].astype(np.int64)
            nb = self.neighbour.astype(np.int64)
            assert np.all(o < nb), "owner < neighbour on internal faces"
            key = o * n + nb
            assert np.all(np.diff(key) > 0), "internal faces in (owner, neighbour) order, no duplicates"
        pos = nif
        for bctype, nfc, sf in self.patches:
            assert bctype in BC_TYPES, bctype
            assert sf == pos, "patches are contiguous and follow the internal faces"
            pos += nfc
        assert pos == nf
        ws, wc = self.wall_range()
        if wc:
            walls = [p for p in self.patches if p[0] in WALL_TYPES]
            assert walls[-1][2] + walls[-1][1] - ws == wc, "wall patches must be contiguous"


# ---------------------------------------------------------------------------
# Spacing helpers
# ---------------------------------------------------------------------------

def uniform_spacing(n: int) -> np.ndarray:
    return np.linspace(0.0, 1.0, n + 1)


def graded_spacing(n: int, ratio: float, max_aspect: float = 100.0) -> np.ndarray:
    """Node coordinates on [0,1] graded geometrically toward 0.

    The first cell is ``(1/n)/max_aspect`` thick; each following cell grows by
    ``ratio`` until the remaining length divided by the remaining cells is no
    larger than the next geometric size, after which the cells are uniform.
    ``ratio <= 1`` gives a uniform grid.
    """
    if ratio <= 1.0:
        return uniform_spacing(n)
    s = (1.0 / n) / max_aspect
    sizes = []
    used = 0.0
    for c in range(n):
        rem_cells = n - c
        rem_len = 1.0 - used
        if rem_len / rem_cells <= s:
            sizes.extend([rem_len / rem_cells] * rem_cells)
            break
        sizes.append(s)
        used += s
        s *= ratio
    x = np.concatenate([[0.0], np.cumsum(sizes)])
    x[-1] = 1.0
    return x


# ---------------------------------------------------------------------------
# Face list canonicalisation
# ---------------------------------------------------------------------------

def _canonical_polymesh(points, n_cells, in_c1, in_c2, in_nodes, bnd_cell, bnd_nodes,
                        bnd_patch, patch_types) -> PolyMesh:
    """Assemble a PolyMesh from unordered internal and boundary faces.

    ``in_nodes`` rows are oriented so the normal points from ``in_c1`` to
    ``in_c2``; rows whose c1 > c2 are flipped.  ``bnd_nodes`` point outwards.
    Face node arrays are (nfaces, k) with a fixed k per call.
    """
    in_c1 = np.asarray(in_c1, dtype=np.int64)
    in_c2 = np.asarray(in_c2, dtype=np.int64)
    swap = in_c1 > in_c2
    own = np.where(swap, in_c2, in_c1)
    nei = np.where(swap, in_c1, in_c2)
    nodes = in_nodes.copy()
    if swap.any():
        # reversing the node order flips the normal; keep node 0 as the first node
        nodes[swap, 1:] = nodes[swap, 1:][:, ::-1]
    order = np.lexsort((nei, own))
    own, nei, nodes = own[order], nei[order], nodes[order]

    b_order = np.lexsort((bnd_cell, bnd_patch))
    bnd_cell = bnd_cell[b_order]
    bnd_nodes = bnd_nodes[b_order]
    bnd_patch = bnd_patch[b_order]

    nif = own.shape[0]
    patches = []
    pos = nif
    for p, t in enumerate(patch_types):
        cnt = int(np.count_nonzero(bnd_patch == p))
        if cnt:
            patches.append((t, cnt, pos))
            pos += cnt

    k_in = nodes.shape[1]
    k_b = bnd_nodes.shape[1]
    nf = nif + bnd_cell.shape[0]
    counts = np.concatenate([np.full(nif, k_in, np.int64), np.full(bnd_cell.shape[0], k_b, np.int64)])
    face_ptr = np.zeros(nf + 1, np.int64)
    np.cumsum(counts, out=face_ptr[1:])
    face_nodes = np.concatenate([nodes.reshape(-1), bnd_nodes.reshape(-1)]).astype(np.int32)
    owner = np.concatenate([own, bnd_cell]).astype(np.int32)
    return PolyMesh(points=np.ascontiguousarray(points, dtype=np.float64),
                    face_ptr=face_ptr.astype(np.int32), face_nodes=face_nodes,
                    owner=owner, neighbour=nei.astype(np.int32), n_cells=int(n_cells),
                    patches=patches)


# Patch ids used by the box generators: x0, x1, y0, y1, z0, z1.
_BOX_SIDES = ("x0", "x1", "y0", "y1", "z0", "z1")


def _patch_types(bc: dict) -> List[str]:
    return [bc.get(s, "symmetry") for s in _BOX_SIDES]


DEFAULT_BC = {"x0": "wall", "x1": "symmetry", "y0": "symmetry", "y1": "symmetry",
              "z0": "symmetry", "z1": "symmetry"}


def _box_points(xs, ys, zs, perturb: float, seed: Optional[int]):
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if perturb > 0.0:
        rng = np.random.default_rng(seed)

        def local(h):
            d = np.diff(h)
            loc = np.empty(len(h))
            loc[0], loc[-1] = d[0], d[-1]
            loc[1:-1] = np.minimum(d[:-1], d[1:])
            return loc

        lx, ly, lz = local(xs), local(ys), local(zs)
        LZ, LY, LX = np.meshgrid(lz, ly, lx, indexing="ij")
        delta = rng.uniform(-perturb, perturb, size=pts.shape)
        delta[:, 0] *= LX.ravel()
        delta[:, 1] *= LY.ravel()
        delta[:, 2] *= LZ.ravel()
        ii = np.arange(nx + 1)
        jj = np.arange(ny + 1)
        kk = np.arange(nz + 1)
        KK, JJ, II = np.meshgrid(kk, jj, ii, indexing="ij")
        interior = ((II > 0) & (II < nx) & (JJ > 0) & (JJ < ny) & (KK > 0) & (KK < nz)).ravel()
        pts[interior] += delta[interior]
    return pts


def hex_box(nx: int, ny: Optional[int] = None, nz: Optional[int] = None, *,
            perturb: float = 0.0, seed: Optional[int] = 0,
            grade_axis: Optional[int] = None, grade_ratio: float = 1.0,
            grade_max_aspect: float = 100.0, bc: Optional[dict] = None) -> PolyMesh:
    """Hexahedral box [0,1]^3 with nx*ny*nz cells in natural (x-fastest) order.

    ``perturb`` moves every interior vertex by U[-perturb, perturb] times the
    smaller adjacent spacing along each axis (seeded).  ``grade_axis`` (0,1,2)
    grades that axis geometrically toward 0 with ``grade_ratio``.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    bc = DEFAULT_BC if bc is None else {**DEFAULT_BC, **bc}
    axes = []
    for ax, n in enumerate((nx, ny, nz)):
        if grade_axis == ax:
            axes.append(graded_spacing(n, grade_ratio, grade_max_aspect))
        else:
            axes.append(uniform_spacing(n))
    xs, ys, zs = axes
    pts = _box_points(xs, ys, zs, perturb, seed)

    px, pxy = nx + 1, (nx + 1) * (ny + 1)

    def P(i, j, k):
        return i + px * j + pxy * k

    def C(i, j, k):
        return i + nx * (j + ny * k)

    c1s, c2s, nds = [], [], []
    # x-faces at x index i+1, between (i,j,k) and (i+1,j,k); normal +x
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx - 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c1s.append(C(i, j, k)); c2s.append(C(i + 1, j, k))
    nds.append(np.stack([P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1)], 1))
    # y-faces, normal +y
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny - 1), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c1s.append(C(i, j, k)); c2s.append(C(i, j + 1, k))
    nds.append(np.stack([P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k)], 1))
    # z-faces, normal +z
    k, j, i = np.meshgrid(np.arange(nz - 1), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c1s.append(C(i, j, k)); c2s.append(C(i, j, k + 1))
    nds.append(np.stack([P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1)], 1))

    bc_cells, bc_nodes, bc_patch = [], [], []
    # x0 (normal -x) / x1 (normal +x)
    k, j = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    j, k = j.ravel(), k.ravel()
    z = np.zeros_like(j)
    bc_cells.append(C(z, j, k))
    bc_nodes.append(np.stack([P(z, j, k), P(z, j, k + 1), P(z, j + 1, k + 1), P(z, j + 1, k)], 1))
    bc_patch.append(np.full(j.shape, 0))
    e = np.full_like(j, nx)
    bc_cells.append(C(e - 1, j, k))
    bc_nodes.append(np.stack([P(e, j, k), P(e, j + 1, k), P(e, j + 1, k + 1), P(e, j, k + 1)], 1))
    bc_patch.append(np.full(j.shape, 1))
    # y0 (normal -y) / y1 (normal +y)
    k, i = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    i, k = i.ravel(), k.ravel()
    z = np.zeros_like(i)
    bc_cells.append(C(i, z, k))
    bc_nodes.append(np.stack([P(i, z, k), P(i + 1, z, k), P(i + 1, z, k + 1), P(i, z, k + 1)], 1))
    bc_patch.append(np.full(i.shape, 2))
    e = np.full_like(i, ny)
    bc_cells.append(C(i, e - 1, k))
    bc_nodes.append(np.stack([P(i, e, k), P(i, e, k + 1), P(i + 1, e, k + 1), P(i + 1, e, k)], 1))
    bc_patch.append(np.full(i.shape, 3))
    # z0 (normal -z) / z1 (normal +z)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i, j = i.ravel(), j.ravel()
    z = np.zeros_like(i)
    bc_cells.append(C(i, j, z))
    bc_nodes.append(np.stack([P(i, j, z), P(i, j + 1, z), P(i + 1, j + 1, z), P(i + 1, j, z)], 1))
    bc_patch.append(np.full(i.shape, 4))
    e = np.full_like(i, nz)
    bc_cells.append(C(i, j, e - 1))
    bc_nodes.append(np.stack([P(i, j, e), P(i + 1, j, e), P(i + 1, j + 1, e), P(i, j + 1, e)], 1))
    bc_patch.append(np.full(i.shape, 5))

    m = _canonical_polymesh(pts, nx * ny * nz, np.concatenate(c1s), np.concatenate(c2s),
                            np.concatenate(nds), np.concatenate(bc_cells), np.concatenate(bc_nodes),
                            np.concatenate(bc_patch), _patch_types(bc))
    return m


def _tri_normal_sign(pts, tri, towards):
    """Sign of (p1-p0)x(p2-p0) . (towards - p0) for each row."""
    p0, p1, p2 = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    return np.einsum("ij,ij->i", nrm, towards - p0)


def tet_box(nx: int, ny: Optional[int] = None, nz: Optional[int] = None, *,
            perturb: float = 0.0, seed: Optional[int] = 0,
            grade_axis: Optional[int] = None, grade_ratio: float = 1.0,
            grade_max_aspect: float = 100.0, permute_seed: Optional[int] = None,
            bc: Optional[dict] = None) -> PolyMesh:
    """Hex box split into 6 tetrahedra per hex (Kuhn split along the (0,0,0)-(1,1,1) diagonal).

    The split is conforming: every hex face is cut along the diagonal from its
    lowest to its highest corner on both sides.  ``permute_seed`` renumbers the
    cells with a seeded random permutation before the faces are put in
    polyMesh order.
    """
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    bc = DEFAULT_BC if bc is None else {**DEFAULT_BC, **bc}
    axes = []
    for ax, n in enumerate((nx, ny, nz)):
        axes.append(graded_spacing(n, grade_ratio, grade_max_aspect) if grade_axis == ax
                    else uniform_spacing(n))
    xs, ys, zs = axes
    pts_unperturbed = _box_points(xs, ys, zs, 0.0, None)
    pts = _box_points(xs, ys, zs, perturb, seed) if perturb > 0 else pts_unperturbed

    px, pxy = nx + 1, (nx + 1) * (ny + 1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel().astype(np.int64), j.ravel().astype(np.int64), k.ravel().astype(np.int64)
    nh = i.shape[0]
    base = i + px * j + pxy * k
    off = {(a, b, c): a + px * b + pxy * c for a in (0, 1) for b in (0, 1) for c in (0, 1)}
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = np.empty((nh, 6, 4), np.int64)
    for t, pm in enumerate(perms):
        v = [0, 0, 0]
        verts = [base + off[tuple(v)]]
        for ax in pm:
            v[ax] = 1
            verts.append(base + off[tuple(v)])
        tets[:, t, :] = np.stack(verts, 1)
    tets = tets.reshape(-1, 4)                       # cell id = 6*hex + t
    n_cells = tets.shape[0]

    # 4 faces per tet, each opposite one vertex
    fl = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]], 0)
    fcell = np.tile(np.arange(n_cells, dtype=np.int64), 4)
    key = np.sort(fl, axis=1)
    nn = pts.shape[0]
    code = (key[:, 0] * nn + key[:, 1]) * nn + key[:, 2]
    order = np.argsort(code, kind="stable")
    code_s = code[order]
    same_next = np.zeros(code_s.shape[0], bool)
    same_next[:-1] = code_s[:-1] == code_s[1:]
    first = np.nonzero(same_next)[0]
    a_idx, b_idx = order[first], order[first + 1]
    paired = np.zeros(code_s.shape[0], bool)
    paired[first] = True
    paired[first + 1] = True
    single = order[~paired]

    cent = pts_unperturbed[tets].mean(axis=1)
    # internal faces: orient from cell(a) to cell(b)
    c1, c2 = fcell[a_idx], fcell[b_idx]
    tri = key[a_idx].copy()
    s = _tri_normal_sign(pts_unperturbed, tri, cent[c2])
    flip = s < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    # boundary faces: outward = away from own centroid
    bcell = fcell[single]
    btri = key[single].copy()
    s = _tri_normal_sign(pts_unperturbed, btri, cent[bcell])
    flip = s > 0
    btri[flip, 1], btri[flip, 2] = btri[flip, 2].copy(), btri[flip, 1].copy()
    fc = pts_unperturbed[btri].mean(axis=1)
    eps = 1e-9
    bpatch = np.full(btri.shape[0], -1)
    bpatch[np.abs(fc[:, 0] - 0.0) < eps] = 0
    bpatch[np.abs(fc[:, 0] - 1.0) < eps] = 1
    bpatch[np.abs(fc[:, 1] - 0.0) < eps] = 2
    bpatch[np.abs(fc[:, 1] - 1.0) < eps] = 3
    bpatch[np.abs(fc[:, 2] - 0.0) < eps] = 4
    bpatch[np.abs(fc[:, 2] - 1.0) < eps] = 5
    assert (bpatch >= 0).all()

    if permute_seed is not None:
        perm = np.random.default_rng(permute_seed).permutation(n_cells)
        c1, c2, bcell = perm[c1], perm[c2], perm[bcell]

    return _canonical_polymesh(pts, n_cells, c1, c2, tri, bcell, btri, bpatch, _patch_types(bc))


# ---------------------------------------------------------------------------
# polyMesh file I/O (the reader's "native" and OpenFOAM ASCII forms)
# ---------------------------------------------------------------------------

_FOAM_BANNER = [
    "/*--------------------------------*- C++ -*----------------------------------*\\",
    "| =========                 |                                                 |",
    "| \\\\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox           |",
    "|  \\\\    /   O peration     | Version:  4.0                                   |",
    "|   \\\\  /    A nd           | Web:      www.OpenFOAM.org                      |",
    "|    \\\\/     M anipulation  |                                                 |",
    "\\*---------------------------------------------------------------------------*/",
    "FoamFile",
    "{",
    "    version     2.0;",
    "    format      ascii;",
    "    class       {cls};",
    "{note}",
    "    location    \"constant/polyMesh\";",
    "    object      {obj};",
    "}",
    "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //",
]


def _foam_header(cls, obj, note):
    lines = [ln.replace("{cls}", cls).replace("{obj}", obj) for ln in _FOAM_BANNER]
    lines[12] = note
    return "\n".join(lines) + "\n"


def write_polymesh(mesh: PolyMesh, directory: str, fmt: str = "native") -> None:
    """Write ``points faces owner neighbour boundary`` into ``directory``.

    ``fmt="native"``: the reference's native layout -- owner starts with
    numNodes, numCells, numFaces, numInnerFaces, indices are 1-based
    (src/mesh_geometry_and_topology.f90:473-480,741-757).
    ``fmt="foam"``: OpenFOAM ASCII with the 17/18-line headers the reference
    skips (src/mesh_geometry_and_topology.f90:488-561); 0-based indices.
    """
    os.makedirs(directory, exist_ok=True)
    nf, nif, n = mesh.n_faces, mesh.n_inner, mesh.n_cells
    fp, fn = mesh.face_ptr, mesh.face_nodes
    counts = np.diff(fp)
    if fmt == "native":
        np.savetxt(os.path.join(directory, "points"), mesh.points, fmt="%.17g")
        with open(os.path.join(directory, "owner"), "w") as f:
            f.write(f"{mesh.points.shape[0]}\n{n}\n{nf}\n{nif}\n")
            np.savetxt(f, mesh.owner.astype(np.int64) + 1, fmt="%d")
        np.savetxt(os.path.join(directory, "neighbour"), mesh.neighbour.astype(np.int64) + 1, fmt="%d")
        with open(os.path.join(directory, "faces"), "w") as f:
            lines = []
            for fi in range(nf):
                nodes = fn[fp[fi]:fp[fi + 1]] + 1
                lines.append(f"{counts[fi]} " + " ".join(map(str, nodes.tolist())))
            f.write("\n".join(lines) + "\n")
    elif fmt == "foam":
        if np.any(counts > 9):
            raise ValueError("the reference's OpenFOAM faces reader handles single-digit vertex counts only")
        note = f"    note        \"nPoints:{mesh.points.shape[0]}  nCells:{n}  nFaces:{nf}  nInternalFaces:{nif}\";"
        with open(os.path.join(directory, "points"), "w") as f:
            f.write(_foam_header("vectorField", "points", "    arch        \"LSB;label=32;scalar=64\";"))
            f.write(f"\n\n{mesh.points.shape[0]}\n(\n")
            f.write("\n".join(f"({p[0]:.17g} {p[1]:.17g} {p[2]:.17g})" for p in mesh.points))
            f.write("\n)\n")
        with open(os.path.join(directory, "owner"), "w") as f:
            f.write(_foam_header("labelList", "owner", note))
            f.write(f"\n\n{nf}\n(\n")
            np.savetxt(f, mesh.owner, fmt="%d")
            f.write(")\n")
        with open(os.path.join(directory, "neighbour"), "w") as f:
            f.write(_foam_header("labelList", "neighbour", note))
            f.write(f"\n\n{nif}\n(\n")
            np.savetxt(f, mesh.neighbour, fmt="%d")
            f.write(")\n")
        with open(os.path.join(directory, "faces"), "w") as f:
            f.write(_foam_header("faceList", "faces", "    arch        \"LSB;label=32;scalar=64\";"))
            f.write(f"\n\n{nf}\n(\n")
            f.write("\n".join(f"{counts[fi]}(" + " ".join(map(str, fn[fp[fi]:fp[fi + 1]].tolist())) + ")"
                              for fi in range(nf)))
            f.write("\n)\n")
    else:
        raise ValueError(fmt)
    with open(os.path.join(directory, "boundary"), "w") as f:
        f.write("# bctype nFaces startFace\n")
        for bctype, nfc, sf in mesh.patches:
            f.write(f"{bctype} {nfc} {sf}\n")


def read_polymesh_native(directory: str) -> PolyMesh:
    """Read the native layout written by :func:`write_polymesh` (fmt="native")."""
    with open(os.path.join(directory, "owner")) as f:
        nn, n, nf, nif = (int(f.readline()) for _ in range(4))
        owner = np.loadtxt(f, dtype=np.int64).reshape(-1) - 1
    neighbour = np.loadtxt(os.path.join(directory, "neighbour"), dtype=np.int64).reshape(-1) - 1
    points = np.loadtxt(os.path.join(directory, "points")).reshape(-1, 3)
    ptr = [0]
    nodes: List[int] = []
    with open(os.path.join(directory, "faces")) as f:
        for line in f:
            v = [int(t) for t in line.split()]
            if not v:
                continue
            nodes.extend(x - 1 for x in v[1:1 + v[0]])
            ptr.append(len(nodes))
    patches = []
    with open(os.path.join(directory, "boundary")) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            t, a, b = s.split()[:3]
            patches.append((t, int(a), int(b)))
    assert points.shape[0] == nn and owner.shape[0] == nf and neighbour.shape[0] == nif
    return PolyMesh(points=points, face_ptr=np.asarray(ptr, np.int32), face_nodes=np.asarray(nodes, np.int32),
                    owner=owner.astype(np.int32), neighbour=neighbour.astype(np.int32), n_cells=n,
                    patches=patches)
