"""Host side of libmgsb on the CPU: .msh loading and surface extraction
against the reference's own mesh files, the box generator, the precompute
and the level bindings against the oracle."""
import os

import numpy as np
import pytest

import oracle as O
from multigridsoftbody_amd import MgsbError, Mesh


def parse_msh(path):
    """Independent reader of gmsh 2.2 ASCII: nodes, tets, type-2 triangles."""
    toks = open(path).read().split()
    i = toks.index("$Nodes") + 1
    n = int(toks[i])
    nodes = np.array(toks[i + 1:i + 1 + 4 * n], float).reshape(n, 4)[:, 1:]
    i = toks.index("$Elements") + 1
    ne = int(toks[i])
    i += 1
    tets, tris = [], []
    for _ in range(ne):
        typ, ntag = int(toks[i + 1]), int(toks[i + 2])
        nn = {2: 3, 4: 4, 15: 1, 1: 2}[typ]
        ids = [int(t) - 1 for t in toks[i + 3 + ntag:i + 3 + ntag + nn]]
        (tets if typ == 4 else tris if typ == 2 else []).append(ids)
        i += 3 + ntag + nn
    return nodes, np.array(tets), np.array(tris)


# counts measured from the reference's data/ meshes
MESHES = [
    ("Y_1_10_1.msh", 55, 125, 84, 80.0),
    ("Y_4_40_4.msh", 817, 2606, 1344, 80.0),
    ("cube10_1_1.msh", None, None, None, None),
    ("cube40_4_4.msh", None, None, None, None),
    ("cube120_12_12.msh", 10937, 48087, 12096, 80.0),
    ("liver700_3500.msh", None, None, None, None),
]


@pytest.mark.parametrize("name,nv,nt,ns,vol", MESHES)
def test_load_msh_matches_file(golden, name, nv, nt, ns, vol):
    path = os.path.join(golden, name)
    m = Mesh.load_msh(path)
    nodes, tets, tris = parse_msh(path)
    assert m.nverts == len(nodes) and m.ntets == len(tets)
    assert np.array_equal(m.tets, tets)
    assert np.allclose(m.pos, nodes.astype(np.float32))
    if nv is not None:
        assert (m.nverts, m.ntets, m.ntris) == (nv, nt, ns)
    # TetFaceExtraction's boundary set == gmsh's own surface triangles
    key = lambda a: set(map(tuple, np.sort(a, axis=1)))  # noqa: E731
    assert key(m.tris) == key(tris)
    # windings point outward: opposite vertex lies behind each face
    P = m.pos.astype(np.float64)
    n = np.cross(P[m.tris[:, 1]] - P[m.tris[:, 0]], P[m.tris[:, 2]] - P[m.tris[:, 0]])
    assert np.all(np.einsum("ij,ij->i", n, P[m.tri_opp] - P[m.tris[:, 0]]) < 0)
    pc = m.precompute(8000.0)
    if vol is not None:
        assert abs(pc["vol"].sum(dtype=np.float64) - vol) < 1e-4


def test_surface_order_is_reference_set_order(golden):
    """std::set<FaceKey> with FaceKey::operator< (SoftObject.cpp:171-178)
    iterates sorted keys in descending order."""
    m = Mesh.load_msh(os.path.join(golden, "Y_1_10_1.msh"))
    keys = [tuple(sorted(t)) for t in m.tris.tolist()]
    assert keys == sorted(keys, reverse=True)


def test_box_generator():
    m = Mesh.box(4, 40, 4)
    assert (m.nverts, m.ntets) == (1025, 3840)
    P = m.pos.astype(np.float64)
    t = m.tets
    det = np.einsum("ij,ij->i", np.cross(P[t[:, 1]] - P[t[:, 0]], P[t[:, 2]] - P[t[:, 0]]), P[t[:, 3]] - P[t[:, 0]])
    assert np.all(det > 0)
    assert np.isclose(det.sum() / 6, 80.0)
    # closed surface: 2 triangles per boundary quad
    assert m.ntris == 2 * 2 * (4 * 40 + 4 * 4 + 40 * 4)
    assert int((np.abs(m.pos[:, 1] - 10) < 1e-5).sum()) == 25
    med = Mesh.box(12, 120, 12)
    assert (med.nverts, med.ntets) == (20449, 103680)


def test_blob_generator():
    m = Mesh.blob(6, 3, 4, 2.5, 0.15, seed=0, h=0.5)
    assert m.ntets > 0 and m.ntets % 6 == 0
    pc = m.precompute(1.0)
    assert np.all(pc["vol"] > 0)
    # a coarser lattice at 3h nests into the same voxel grid
    c = Mesh.blob(6, 3, 4, 2.5, 0.15, seed=0, h=1.5)
    assert np.allclose(np.round(c.pos / 0.5), c.pos / 0.5, atol=1e-4)


@pytest.mark.parametrize("name", ["Y_1_10_1.msh", "Y_4_40_4.msh", "liver700_3500.msh"])
def test_precompute_matches_oracle_bitwise(golden, name):
    m = Mesh.load_msh(os.path.join(golden, name))
    pc = m.precompute(8000.0, 1.0)
    L = O.Level(m.pos, m.tets, m.tris, stiffness=8000.0)
    for k, sel in [("B", O.B), ("AC", O.AC), ("vol", O.VOL), ("mass", O.MASS), ("diag", O.DIAG)]:
        assert np.array_equal(pc[k].reshape(-1).astype(np.float64), L.get(sel)), k


def test_degenerate_tet_is_an_error_not_an_exit():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    m = Mesh(pos, np.array([[0, 1, 2, 3]]))
    with pytest.raises(MgsbError, match="degenerate"):
        m.precompute(1.0)


@pytest.mark.parametrize("pair", [("Y_4_40_4.msh", "Y_1_10_1.msh"), ("cube40_4_4.msh", "cube10_1_1.msh"),
                                  ("cube120_12_12.msh", "cube40_4_4.msh")])
def test_bindings_match_brute_force_oracle(golden, pair):
    """Spatial-hash binding == PDSolver_MG's brute-force scans, bit for bit."""
    fine = Mesh.load_msh(os.path.join(golden, pair[0]))
    coarse = Mesh.load_msh(os.path.join(golden, pair[1]))
    Lf = O.Level(fine.pos, fine.tets, fine.tris)
    Lc = O.Level(coarse.pos, coarse.tets, coarse.tris)
    mg = O.TwoLevel(Lc, Lf)
    iids, iw, aids, aw = mg.maps()
    unf_f, unf_c = mg.unfound()
    ids0, w0, u0 = coarse.bind(fine.pos, 0)
    ids1, w1, u1 = fine.bind(coarse.pos, 1)
    assert u0 == unf_f and u1 == unf_c
    assert np.array_equal(ids0, iids) and np.array_equal(w0, iw)
    assert np.array_equal(ids1, aids) and np.array_equal(w1, aw)


def test_nested_box_bindings():
    f = Mesh.box(12, 120, 12)
    c = Mesh.box(4, 40, 4)
    ids, w, unf = c.bind(f.pos, 0)
    assert unf == 0
    assert np.allclose(w.sum(1), 1, atol=1e-5)
    rec = (w[:, :, None] * c.pos[ids].astype(np.float64)).sum(1)
    assert np.allclose(rec, f.pos, atol=1e-5)


def _tile_layout_checks(mesh):
    """The tiled local step's layout against its definition: tiles of 256
    consecutive tets; each tile's vertices are the distinct corners of its
    tets; every (tile, vertex) partial has one slot, the slots of a vertex
    are contiguous in ascending tile order; a tile's corner array holds each
    of its corners once, in the row of the corner's vertex, in (tet, corner)
    order, within the 1024 entries of the kernel's LDS arrays; and summing
    corner values through the layout gives every vertex the sum over its
    incident corners."""
    L = mesh.tiles()
    tets = np.asarray(mesh.tets).reshape(-1, 4)
    T, V = tets.shape[0], mesh.nverts
    prp, tvo, tvs, tlr, pos = L["prp"], L["tvo"], L["tvs"], L["tlr"], L["pos"]
    nt = (T + 255) // 256
    npart = int(tvo[-1])
    assert len(tvo) == nt + 1 and tvo[0] == 0 and np.all(np.diff(tvo) > 0)
    assert prp[0] == 0 and prp[-1] == npart and np.all(np.diff(prp) >= 0)
    assert np.array_equal(np.sort(tvs), np.arange(npart))
    assert pos.min() >= 0 and pos.max() < 1024
    owner = np.searchsorted(prp, tvs, side="right") - 1  # vertex of each partial slot
    tile_of_slot = np.empty(npart, np.int64)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((T, 4))
    part = np.zeros(npart)
    for k in range(nt):
        t0, t1 = 256 * k, min(T, 256 * k + 256)
        vk = owner[tvo[k]:tvo[k + 1]]
        assert np.array_equal(vk, np.unique(tets[t0:t1])), k
        tile_of_slot[tvs[tvo[k]:tvo[k + 1]]] = k
        rows = tlr[tvo[k] + k:tvo[k + 1] + k + 1]
        assert rows[0] == 0 and rows[-1] == 4 * (t1 - t0) and np.all(np.diff(rows) > 0)
        p = pos[t0:t1].ravel()
        assert np.array_equal(np.sort(p), np.arange(4 * (t1 - t0)))
        loc = np.searchsorted(vk, tets[t0:t1].ravel())
        assert np.all(rows[loc] <= p) and np.all(p < rows[loc + 1])
        order = np.argsort(p)
        assert np.all(np.diff(loc[order]) >= 0)
        for l in range(len(vk)):  # (tet, corner) order inside each row
            assert np.all(np.diff(order[rows[l]:rows[l + 1]]) > 0)
        lds = np.zeros(4 * (t1 - t0))
        lds[p] = f[t0:t1].ravel()
        part[tvs[tvo[k]:tvo[k + 1]]] = np.add.reduceat(lds, rows[:-1])
    for v in range(0, V, max(1, V // 200)):
        assert np.all(np.diff(tile_of_slot[prp[v]:prp[v + 1]]) > 0)
    want = np.zeros(V)
    np.add.at(want, tets.ravel(), f.ravel())
    got = np.add.reduceat(np.append(part, 0.0), prp[:-1]) * (np.diff(prp) > 0)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    return npart


@pytest.mark.parametrize("which", ["box4", "box12", "liver", "blob"])
def test_tile_layout(golden, which):
    from multigridsoftbody_amd import workloads as W

    mesh = {"box4": lambda: W.beam(4), "box12": lambda: W.beam(12),
            "liver": lambda: Mesh.load_msh(os.path.join(golden, "liver700_3500.msh")),
            "blob": lambda: Mesh.blob(h=0.45)}[which]()
    npart = _tile_layout_checks(mesh)
    print(f"[tiles {which}] {mesh.ntets} tets, {npart} partials: {npart / mesh.nverts:.2f} per vertex, "
          f"{4 * mesh.ntets / npart:.1f} corners per partial")
