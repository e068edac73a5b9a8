"""Synthetic polycrystal specimens in icFlow's data model.

The reference reads gmsh meshes with per-tet grain tags and inserts cohesive zones with
CZInsertionTool; none of its scenario files ship with it.  This module builds equivalent
meshes from scratch -- a hex grid split into 6 tets per hex, grains from seeded Voronoi
seeds -- and then follows the reference's conventions exactly where they decide data layout:

  * tet face f has vertices tet[CONV[f]] (CZInsertionTool.cs:183-187, myConvention);
  * a CZ joins the lower-numbered tet e0 and the higher-numbered tet e1 of an inter-grain
    face; vrts[0..2] are e0's face vertices, vrts[3..5] the matching vertices of e1, with
    the reference's rotation and 4<->5 swap (ExtendedCZ, CZInsertionTool.cs:199-233);
  * every node on an inter-grain face is split into one copy per grain of its incident tets,
    grains in order of first appearance over the node's tets; the last grain keeps the
    original node and the copies follow it in reverse order (ExtendedNode.SplitNode,
    CZInsertionTool.cs:22-35, AddAfter);
  * faces after insertion: exposed faces first, then per CZ the e0 face and the e1 face
    (CZInsertionTool.cs:448-485);
  * surface elements = tets with a node on any face; adjacent faces = union over the tet's
    vertices of each vertex's faces, in first-insertion order (Mesh.cs:665-678, Element.cs:15-20).

Everything is vectorised numpy so that the 1.3M-tet specimen builds in seconds.
"""
from dataclasses import dataclass, field

import numpy as np

CONV = np.array([[3, 1, 2], [0, 3, 2], [0, 1, 3], [1, 0, 2]], dtype=np.int64)


def jdet(x):
    """Jdet of element_kernels2.h:47 for tets x[..., 4, 3] (6 x signed volume)."""
    x1, x2, x3, x4 = x[..., 0, :], x[..., 1, :], x[..., 2, :], x[..., 3, :]
    d21, d32, d43 = x2 - x1, x3 - x2, x4 - x3
    d23, d34, d12 = x2 - x3, x3 - x4, x1 - x2
    return (d21[..., 0] * (d23[..., 1] * d34[..., 2] - d34[..., 1] * d23[..., 2])
            + d32[..., 0] * (d34[..., 1] * d12[..., 2] - d12[..., 1] * d34[..., 2])
            + d43[..., 0] * (d12[..., 1] * d23[..., 2] - d23[..., 1] * d12[..., 2]))


def hex_grid_tets(nx, ny, nz, size, origin=(0.0, 0.0, 0.0)):
    """Nodes (i fastest) and 6 Kuhn tets per hex, all with Jdet > 0."""
    lx, ly, lz = size
    xs = origin[0] + np.linspace(0.0, lx, nx + 1)
    ys = origin[1] + np.linspace(0.0, ly, ny + 1)
    zs = origin[2] + np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    x0 = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    e = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for a, b, _ in perms:
        p1 = e[a]
        p2 = (e[a][0] + e[b][0], e[a][1] + e[b][1], e[a][2] + e[b][2])
        tets.append(np.stack([nid(i, j, k), nid(i + p1[0], j + p1[1], k + p1[2]),
                              nid(i + p2[0], j + p2[1], k + p2[2]), nid(i + 1, j + 1, k + 1)], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    neg = jdet(x0[tets]) < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    assert (jdet(x0[tets]) > 0).all()
    return x0, tets.astype(np.int64)


def voronoi_grains(x0, tets, n_grains, seed, box_lo, box_hi):
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(box_lo, box_hi, size=(n_grains, 3))
    cent = x0[tets].mean(axis=1)
    _, g = cKDTree(seeds).query(cent)
    return g.astype(np.int64)


@dataclass
class Mesh:
    """One mesh of a MeshCollection (Mesh.cs): nodes, tets, faces, CZs."""
    x0: np.ndarray                 # (n, 3)
    elems: np.ndarray              # (T, 4) local node ids
    faces: np.ndarray              # (F, 3) local node ids
    exposed: np.ndarray            # (F,) bool
    czs: np.ndarray                # (C, 6) local node ids
    cz_faces: np.ndarray           # (C, 2) local face ids
    deformable: bool = True
    grain: np.ndarray = None


def insert_cohesive_zones(x0, tets, grain):
    """CZInsertionTool.InsertCohesiveElements restated on arrays (CZInsertionTool.cs:371-491)."""
    T = len(tets)
    ft = tets[:, CONV].reshape(-1, 3)                 # (4T, 3) face vertices, tet-major
    key = np.sort(ft, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    owner = np.repeat(np.arange(T), 4)
    which = np.tile(np.arange(4), T)
    # per unique face: first and second (tet, which) in tet order (Mesh.IdentifyParentsOfTriangles)
    order = np.argsort(inv, kind="stable")
    first = np.full(len(cnt), -1, np.int64)
    second = np.full(len(cnt), -1, np.int64)
    uinv = inv[order]
    start = np.r_[0, np.flatnonzero(np.diff(uinv)) + 1]
    first[uinv[start]] = order[start]
    has2 = cnt[uinv[start]] == 2
    second[uinv[start[has2]]] = order[start[has2] + 1]
    # the mesh's face list: boundary faces, then inter-grain faces, each in order of the
    # (tet, local face) that first holds them
    bnd = np.flatnonzero(cnt == 1)
    bnd = bnd[np.argsort(first[bnd])]
    inner = np.flatnonzero(cnt == 2)
    inner = inner[grain[owner[first[inner]]] != grain[owner[second[inner]]]]
    inner = inner[np.argsort(first[inner])]

    e0, f0 = owner[first[inner]], which[first[inner]]
    e1, f1 = owner[second[inner]], which[second[inner]]
    # relative orientation (ExtendedCZ ctor): e1 face vertex r coincides with e0 face vertex 0
    v0 = tets[e0, CONV[f0, 0]]
    e1v = tets[e1[:, None], CONV[f1]]                 # (C, 3)
    rel = np.argmax(e1v == v0[:, None], axis=1)
    assert (e1v[np.arange(len(rel)), rel] == v0).all()

    # node splitting
    nn = len(x0)
    on_cz = np.zeros(nn, bool)
    on_cz[ft[first[inner]].ravel()] = True
    inc_node = tets.ravel()
    inc_tet = np.repeat(np.arange(T), 4)
    inc_grain = grain[inc_tet]
    o = np.lexsort((inc_tet, inc_node))               # by node, then tet
    sn, sg = inc_node[o], inc_grain[o]
    # first appearance of each (node, grain)
    pair = sn * (grain.max() + 1) + sg
    _, fpos = np.unique(pair, return_index=True)
    fpos = np.sort(fpos)                              # node-major, first-appearance order
    pn, pg = sn[fpos], sg[fpos]
    m = np.bincount(pn, minlength=nn)                 # grains per node
    rank = np.arange(len(pn)) - np.r_[0, np.cumsum(m)[:-1]][pn]
    copies = np.where(on_cz, m, 1)
    base = np.r_[0, np.cumsum(copies)[:-1]]
    # new id of (node, grain): last grain keeps the original (offset 0), grain i -> m-1-i
    off = np.where(on_cz[pn], np.where(rank == m[pn] - 1, 0, m[pn] - 1 - rank), 0)
    newid_pairs = base[pn] + off
    lut = dict()
    keyp = pn * (grain.max() + 1) + pg
    srt = np.argsort(keyp)
    keyp_s, newid_s = keyp[srt], newid_pairs[srt]
    tkey = tets * (grain.max() + 1) + grain[:, None]
    new_tets = newid_s[np.searchsorted(keyp_s, tkey)]
    new_x0 = np.repeat(x0, copies, axis=0)
    del lut

    # CZ vertices (ReinitializeVerticeArrays, CZInsertionTool.cs:221-232)
    a = new_tets[e0[:, None], CONV[f0]]
    bidx = (np.arange(3)[None, :] + rel[:, None]) % 3
    b = new_tets[e1[:, None], CONV[f1][np.arange(len(rel))[:, None], bidx]]
    czs = np.concatenate([a, b[:, [0, 2, 1]]], axis=1)

    # faces: exposed first (surfaceFaces order), then e0 / e1 face per CZ
    fb = new_tets[owner[first[bnd]][:, None], CONV[which[first[bnd]]]]
    fc0 = new_tets[e0[:, None], CONV[f0]]
    fc1 = new_tets[e1[:, None], CONV[f1]]
    C = len(inner)
    faces = np.concatenate([fb, np.stack([fc0, fc1], axis=1).reshape(-1, 3)], axis=0)
    exposed = np.r_[np.ones(len(fb), bool), np.zeros(2 * C, bool)]
    cz_faces = len(fb) + np.arange(2 * C).reshape(C, 2)
    return new_x0, new_tets, czs, faces, exposed, cz_faces


def boundary_faces(tets):
    T = len(tets)
    ft = tets[:, CONV].reshape(-1, 3)
    _, inv, cnt = np.unique(np.sort(ft, axis=1), axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    sel = cnt[inv] == 1
    return ft[np.flatnonzero(sel)]


@dataclass
class Collection:
    """MeshCollection (MeshCollection.cs) flattened to global arrays."""
    meshes: list
    x0: np.ndarray = None          # (Nn, 3) all nodes
    elastic: np.ndarray = None     # (T, 4) global ids of deformable tets
    czs: np.ndarray = None         # (C, 6) global ids
    faces: np.ndarray = None       # (F, 3) global ids
    exposed: np.ndarray = None
    anchored: np.ndarray = None    # (Nn,) bool
    alt: np.ndarray = None         # (Nn,) altId or -1
    surf: np.ndarray = None        # (S, 4) global ids of surface tets (all meshes)
    surf_deformable: np.ndarray = None
    adj: list = field(default_factory=list)
    adj_ptr: np.ndarray = None     # CSR of adjacent global face ids per surface tet
    adj_faces: np.ndarray = None
    node_mesh: np.ndarray = None

    def finalize(self):
        offs, xs, el, cz, fc, ex, sf, sdef, nm = 0, [], [], [], [], [], [], [], []
        foffs = 0
        for mi, m in enumerate(self.meshes):
            xs.append(m.x0)
            nm.append(np.full(len(m.x0), mi))
            if m.deformable:
                el.append(m.elems + offs)
                if len(m.czs):
                    cz.append(m.czs + offs)
            fc.append(m.faces + offs)
            ex.append(m.exposed)
            # surface elements: tets with any node on any face of the mesh (Mesh.cs:665-672)
            isn = np.zeros(len(m.x0), bool)
            isn[m.faces.ravel()] = True
            s = np.flatnonzero(isn[m.elems].any(axis=1))
            sf.append(m.elems[s] + offs)
            sdef.append(np.full(len(s), m.deformable))
            offs += len(m.x0)
            foffs += len(m.faces)
        self.x0 = np.concatenate(xs)
        self.node_mesh = np.concatenate(nm)
        self.elastic = np.concatenate(el) if el else np.zeros((0, 4), np.int64)
        self.czs = np.concatenate(cz) if cz else np.zeros((0, 6), np.int64)
        self.faces = np.concatenate(fc)
        self.exposed = np.concatenate(ex)
        self.surf = np.concatenate(sf)
        self.surf_deformable = np.concatenate(sdef)
        if self.anchored is None:
            self.anchored = np.zeros(len(self.x0), bool)
        for mi, m in enumerate(self.meshes):       # non-deformables are anchored (MeshCollection.cs:51-52)
            if not m.deformable:
                self.anchored[self.node_mesh == mi] = True
        self.alt = np.full(len(self.x0), -1, np.int64)
        act = np.flatnonzero(~self.anchored)
        self.alt[act] = np.arange(len(act))
        self._adjacency()
        return self

    def _adjacency(self):
        """Element.FindAdjFaces over Node.faces (Mesh.ConnectFaces): first-insertion order."""
        F = len(self.faces)
        nf_node = self.faces.ravel()
        nf_face = np.repeat(np.arange(F), 3)
        o = np.lexsort((nf_face, nf_node))
        n_sorted, f_sorted = nf_node[o], nf_face[o]
        nptr = np.r_[0, np.cumsum(np.bincount(n_sorted, minlength=len(self.x0)))]
        S = len(self.surf)
        # candidate list per surface tet: vertex-major, then the vertex's faces in face order
        cnts = nptr[self.surf + 1] - nptr[self.surf]             # (S, 4)
        tot = cnts.sum(axis=1)
        rows = np.repeat(np.arange(S), tot)
        vtx = np.repeat(self.surf.ravel(), cnts.ravel())
        within = np.arange(len(vtx)) - np.repeat(np.r_[0, np.cumsum(cnts.ravel())[:-1]], cnts.ravel())
        cand = f_sorted[nptr[vtx] + within]
        seq = np.arange(len(cand))
        # keep first occurrence of each (tet, face)
        k = rows * F + cand
        _, firstpos = np.unique(k, return_index=True)
        keep = np.sort(firstpos)
        rows_k, cand_k = rows[keep], cand[keep]
        self.adj_ptr = np.r_[0, np.cumsum(np.bincount(rows_k, minlength=S))]
        self.adj_faces = cand_k
        del seq


def make_specimen(n, size, n_grains, seed, origin=(0.0, 0.0, 0.0), platens=None, platen_gap=0.0):
    """A hex-grid polycrystal with CZs on every inter-grain face.

    n = (nx, ny, nz) cells, size = (lx, ly, lz) metres.  platens: None or (px, py) cells of
    two rigid one-cell-thick tet slabs above and below the specimen (non-deformable meshes).
    """
    x0, tets = hex_grid_tets(*n, size, origin)
    lo = np.asarray(origin, float)
    hi = lo + np.asarray(size, float)
    if n_grains > 1:
        grain = voronoi_grains(x0, tets, n_grains, seed, lo, hi)
        nx0, ntets, czs, faces, exposed, cz_faces = insert_cohesive_zones(x0, tets, grain)
    else:
        grain = np.zeros(len(tets), np.int64)
        nx0, ntets = x0, tets
        faces = boundary_faces(tets)
        exposed = np.ones(len(faces), bool)
        czs = np.zeros((0, 6), np.int64)
        cz_faces = np.zeros((0, 2), np.int64)
    meshes = [Mesh(nx0, ntets, faces, exposed, czs, cz_faces, True, grain)]
    if platens is not None:
        px, py = platens
        hz = size[2] / n[2]
        for zc in (lo[2] - hz - platen_gap, hi[2] + platen_gap):
            pxs, pt = hex_grid_tets(px, py, 1, (size[0], size[1], hz), (lo[0], lo[1], zc))
            pf = boundary_faces(pt)
            meshes.append(Mesh(pxs, pt, pf, np.ones(len(pf), bool), np.zeros((0, 6), np.int64),
                               np.zeros((0, 2), np.int64), False, None))
    return Collection(meshes)


@dataclass
class State:
    """Per-node tentative kinematics (Node.cs:17-22) and CZ history (CZ.cs)."""
    un: np.ndarray
    vn: np.ndarray
    an: np.ndarray
    xc: np.ndarray
    pmax: np.ndarray   # (C, 3)
    tmax: np.ndarray   # (C, 3)


def synthetic_state(col, consts, seed, strain=1e-4, noise_decades=(-2.5, -0.5), hist=0.8, equal_gp=False,
                    platen_dz=0.0):
    """Seeded kinematic state: affine strain plus per-node Gaussian noise whose scale is
    delta_n * 10^U(noise_decades) (log-uniform, so that CZ openings -- differences of the
    noise of coincident split nodes -- span contact, loading, unloading and failure), random
    velocities / accelerations and CZ history pmax ~ U(0, hist delta_n), tmax ~ U(0, hist
    delta_t).  Rigid meshes are translated by platen_dz along z towards the specimen."""
    rng = np.random.default_rng(seed)
    Nn = len(col.x0)
    deln, delt = consts.deln, consts.delt
    cen = col.x0.mean(axis=0)
    un = strain * (col.x0 - cen) * np.array([1.0, 1.0, -1.0])
    scale = deln * 10.0 ** rng.uniform(noise_decades[0], noise_decades[1], size=(Nn, 1))
    un += rng.normal(0.0, 1.0, size=(Nn, 3)) * scale
    for mi, m in enumerate(col.meshes):
        if not m.deformable:
            sel = col.node_mesh == mi
            zmid = col.x0[sel, 2].mean()
            un[sel] = 0.0
            un[sel, 2] = -np.sign(zmid - cen[2]) * platen_dz
    vn = rng.normal(0.0, 1e-3, size=(Nn, 3))
    an = rng.normal(0.0, 1.0, size=(Nn, 3))
    C = len(col.czs)
    if equal_gp:
        pm = rng.uniform(0, hist * deln, size=(C, 1)).repeat(3, axis=1)
        tm = rng.uniform(0, hist * delt, size=(C, 1)).repeat(3, axis=1)
    else:
        pm = rng.uniform(0, hist * deln, size=(C, 3))
        tm = rng.uniform(0, hist * delt, size=(C, 3))
    # a fraction of CZs starts undamaged (pmax = tmax = 0 branches)
    z = rng.uniform(size=C) < 0.2
    pm[z] = 0.0
    tm[z] = 0.0
    return State(un, vn, an, col.x0 + un, pm, tm)


def batch(collections):
    """Many independent specimens as one collection: their meshes side by side, node, tet,
    CZ and face ids offset per specimen, so that one pass over the flattened unit lists
    assembles every specimen into its own diagonal block of one BCSR matrix (rows are
    numbered specimen by specimen).  Specimen k's nodes are col.node_mesh == k when every
    specimen is one deformable mesh."""
    meshes = [m for c in collections for m in c.meshes]
    anchored = None
    if any(c.anchored is not None for c in collections):
        anchored = np.concatenate([c.anchored if c.anchored is not None else
                                   np.zeros(sum(len(m.x0) for m in c.meshes), bool) for c in collections])
    out = Collection(meshes)
    out.anchored = anchored
    return out.finalize()


def specimen_offsets(col):
    """Unit offsets per specimen of a batched collection whose meshes are its specimens: tets,
    CZs and active rows (n + 1 int32 each, from 0).  Units are mesh-ordered (finalize), so each
    specimen's units are one run."""
    n = len(col.meshes)
    nm = col.node_mesh

    def off(first_nodes):
        k = nm[first_nodes] if len(first_nodes) else np.zeros(0, np.int64)
        assert np.all(np.diff(k) >= 0), "units are not ordered by specimen"
        return np.r_[0, np.cumsum(np.bincount(k, minlength=n))].astype(np.int32)

    tet = off(col.elastic[:, 0])
    cz = off(col.czs[:, 0]) if len(col.czs) else np.zeros(n + 1, np.int32)
    row = off(np.flatnonzero(~col.anchored))
    return tet, cz, row


def lognormal_cubes(n_specimens, seed, median_tets=12000, min_tets=2000, max_tets=40000, sigma=0.6,
                    grains=(8, 64), size=0.02):
    """The batched workload: cubes whose tet counts are lognormal (median `median_tets`,
    log-standard deviation `sigma`, truncated to [min_tets, max_tets]), each a hex grid of
    n^3 cells x 6 tets with n = round((tets/6)^(1/3)) and its own Voronoi grain count drawn
    uniformly from `grains` and its own grain seed.  Specimens are laid out along x, apart."""
    return batch(lognormal_specimens(n_specimens, seed, median_tets, min_tets, max_tets, sigma, grains, size))


def lognormal_specimens(n_specimens, seed, median_tets=12000, min_tets=2000, max_tets=40000, sigma=0.6,
                        grains=(8, 64), size=0.02):
    """The specimens of lognormal_cubes as separate (unfinalized) collections."""
    rng = np.random.default_rng(seed)
    cols = []
    x = 0.0
    for k in range(n_specimens):
        t = float(np.clip(rng.lognormal(np.log(median_tets), sigma), min_tets, max_tets))
        n = max(2, int(round((t / 6.0) ** (1.0 / 3.0))))
        g = int(rng.integers(grains[0], grains[1] + 1))
        c = make_specimen((n, n, n), (size, size, size), g, int(rng.integers(1 << 30)), origin=(x, 0.0, 0.0))
        cols.append(c)
        x += 2 * size
    return cols
