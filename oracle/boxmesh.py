"""Box meshes and their continuous node numbering, restated in pure Python for small
cases -- test infrastructure only (tests/ import it; the product never does).

The reference builds a box of elements from vertex tables, matches the nodes of
every face shared by two elements, and labels each node with the smallest 1 + l
reachable through those matches.  This module does the same from the tables,
with no lattice formula, so it checks the product's closed-form lattice ids:

  * elements and vertices: libs/mesh/meshSetupBoxTri2D.cpp:80-113,
    meshSetupBoxQuad2D.cpp:79-106, meshSetupBoxTet3D.cpp:87-189,
    meshSetupBoxHex3D.cpp:87-130 (one rank, non-periodic: i1 = i+1 < NnX);
  * faces: meshSetElementType.cpp:30-90 (face vertex tables);
  * reference nodes: equispaced simplex nodes meshBasisTri2D.cpp:81-92 and
    meshBasisTet3D.cpp:81-94; warp and blend (meshBasisTri2D.cpp:34-40,
    meshBasisTet3D.cpp:34-40) keeps every node on its vertex, edge or face and
    is symmetric, so face matches pair the same nodes; tensor nodes i + j*Nq
    (+ k*Nq^2) of meshBasisQuad2D.cpp:34-47 and meshBasisHex3D.cpp:34-48 are
    placed at i/N, a monotone stand-in for the GLL points with the same matches;
  * face-node matching and the numbering fixed point: meshConnect.cpp,
    meshConnectFaceNodes.cpp, meshConnectNodes.cpp:43-81.

Positions are exact: vertex lattice coordinates times N, barycentric weights
times N, all integers.
"""
from itertools import product

TRI, QUAD, TET, HEX = 3, 4, 6, 12
NAMES = {"Tri": TRI, "Quad": QUAD, "Tet": TET, "Hex": HEX}

# face vertex lists (meshSetElementType.cpp)
FACES = {
    TRI: [(0, 1), (1, 2), (2, 0)],
    QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)],
    TET: [(0, 1, 2), (0, 3, 1), (1, 3, 2), (0, 2, 3)],
    HEX: [(0, 1, 2, 3), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (0, 3, 7, 4), (4, 7, 6, 5)],
}


def _cells(nx, ny, nz, dim):
    if dim == 2:
        return [(i, j, 0) for j in range(ny) for i in range(nx)]
    return [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)]


def elements(etype, nx, ny, nz=1):
    """Vertex coordinates (x, y[, z]) of every element, in the reference's element order."""
    out = []
    if etype == TRI:
        for (i, j, _) in _cells(nx, ny, nz, 2):
            out.append([(i, j), (i + 1, j), (i + 1, j + 1)])
            out.append([(i, j), (i + 1, j + 1), (i, j + 1)])
    elif etype == QUAD:
        for (i, j, _) in _cells(nx, ny, nz, 2):
            out.append([(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)])
    elif etype == TET:
        for (i, j, k) in _cells(nx, ny, nz, 3):
            c = {}
            for (b0, b1, b2) in product((0, 1), repeat=3):
                c[b0 + 2 * b1 + 4 * b2] = (i + b0, j + b1, k + b2)
            for t in ((0, 3, 2, 7), (0, 1, 3, 7), (0, 2, 6, 7), (0, 6, 4, 7), (0, 5, 1, 7), (0, 4, 5, 7)):
                out.append([c[v] for v in t])
    elif etype == HEX:
        for (i, j, k) in _cells(nx, ny, nz, 3):
            out.append([(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                        (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)])
    else:
        raise ValueError("unknown element type")
    return out


def reference_nodes(etype, N):
    """Per node, its weights times N on the element's vertices, in the reference's node order."""
    w = []
    if etype == TRI:
        for n in range(N + 1):
            for m in range(N + 1 - n):
                w.append((N - m - n, m, n))
    elif etype == TET:
        for k in range(N + 1):
            for n in range(N + 1 - k):
                for m in range(N + 1 - n - k):
                    w.append((N - m - n - k, m, n, k))
    elif etype == QUAD:
        # bilinear weights times N^2, scaled back by N below
        for j in range(N + 1):
            for i in range(N + 1):
                w.append(((N - i) * (N - j), i * (N - j), i * j, (N - i) * j))
    elif etype == HEX:
        for k in range(N + 1):
            for j in range(N + 1):
                for i in range(N + 1):
                    a, b = (N - i, i), (N - j, j)
                    c = (N - k, k)
                    w.append((a[0] * b[0] * c[0], a[1] * b[0] * c[0], a[1] * b[1] * c[0], a[0] * b[1] * c[0],
                              a[0] * b[0] * c[1], a[1] * b[0] * c[1], a[1] * b[1] * c[1], a[0] * b[1] * c[1]))
    return w


def positions(etype, nx, ny, nz, N):
    """Exact position (integer tuple, in units of 1/N of an element side for simplices and
    1/N^2 for tensor elements) of every local node l = e*Np + n."""
    els = elements(etype, nx, ny, nz)
    w = reference_nodes(etype, N)
    out = []
    for verts in els:
        d = len(verts[0])
        for ww in w:
            out.append(tuple(sum(wv * v[a] for wv, v in zip(ww, verts)) for a in range(d)))
    return out


def connect_min_ids(etype, nx, ny, nz, N):
    """1 + smallest local index joined to each node through shared-face matches, iterated to a
    fixed point (meshConnectNodes.cpp:54-81)."""
    els = elements(etype, nx, ny, nz)
    w = reference_nodes(etype, N)
    Np = len(w)
    pos = positions(etype, nx, ny, nz, N)
    faces = FACES[etype]
    # nodes on face f: zero weight on every vertex outside f
    on_face = []
    for f in faces:
        outside = [v for v in range(len(els[0])) if v not in f]
        on_face.append([n for n in range(Np) if all(w[n][v] == 0 for v in outside)])
    owners = {}
    for e, verts in enumerate(els):
        for fi, f in enumerate(faces):
            owners.setdefault(frozenset(verts[v] for v in f), []).append((e, fi))
    pairs = []
    for key, own in owners.items():
        if len(own) > 2:
            raise AssertionError("face shared by more than two elements")
        if len(own) < 2:
            continue
        (eM, fM), (eP, fP) = own
        where = {pos[eP * Np + n]: eP * Np + n for n in on_face[fP]}
        for n in on_face[fM]:
            lM = eM * Np + n
            lP = where[pos[lM]]  # KeyError: the face nodes do not coincide
            pairs.append((lM, lP))
            pairs.append((lP, lM))
    ids = [1 + l for l in range(len(els) * Np)]
    changed = True
    while changed:
        changed = False
        for (m, p) in pairs:
            if ids[p] < ids[m]:
                ids[m] = ids[p]
                changed = True
    return ids


def boundary_face_count(etype, nx, ny, nz=1):
    """Faces owned by one element only."""
    els = elements(etype, nx, ny, nz)
    count = {}
    for verts in els:
        for f in FACES[etype]:
            key = frozenset(verts[v] for v in f)
            count[key] = count.get(key, 0) + 1
    return sum(1 for c in count.values() if c == 1)


def signed_measure(verts):
    """Twice the signed area (triangle) or six times the signed volume (tetrahedron)."""
    if len(verts) == 3:
        (x0, y0), (x1, y1), (x2, y2) = verts
        return (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    a, b, c, d = verts
    u = [b[i] - a[i] for i in range(3)]
    v = [c[i] - a[i] for i in range(3)]
    t = [d[i] - a[i] for i in range(3)]
    return (u[0] * (v[1] * t[2] - v[2] * t[1]) - u[1] * (v[0] * t[2] - v[2] * t[0])
            + u[2] * (v[0] * t[1] - v[1] * t[0]))
