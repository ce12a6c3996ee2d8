"""Chain builder for unstructured 2D meshes of triangles and quadrilaterals.

The reference reads chains from files (iqloops.dat, qloops.dat, ncolors.dat;
io.cpp:199-268) written by a mesh generator it does not ship.  This module
builds them from nodes and cells alone, in the layout its solver expects:

  - a chain enters each cell it visits through one face and leaves through
    another; face k of a chain has the chain's cell k on its left and cell
    k-1 on its right; an open chain starts and ends on boundary faces, a
    closed chain lists its first face again at the end (computeRHS.cpp:53-67,
    75-95; ADI.cpp:85-191);
  - every face lies on exactly one chain, or is an orphan face;
  - chains are grouped by colour, the chains of one colour share no cell
    (ADI's chain-by-chain Gauss-Seidel order then equals one solve per
    colour; ncolors.dat holds the counts);
  - cell 0 sits at the start of an open chain or inside a closed one, never
    where the reference's -0 == 0 ghost test misfires (computeRHS.cpp:158,170).

A chain uses two faces of every cell it visits, so a cell can lie on chains
only through an even number of its faces; a triangle always leaves an odd
number over.  The faces left over are orphan faces: interior faces the solver
treats one by one (first-order flux, Jacobians on the ADI diagonal).  Choosing
them is a T-join of the triangles in the cell adjacency graph: an interior
face set in which every triangle has an odd and every quadrilateral an even
number of faces.  Adjacent triangles are paired greedily first (one orphan
face per pair); triangles left unpaired are joined along shortest paths of
cells.  Each component of the cell graph needs an even number of triangles.
Quadrilaterals pair opposite faces; a cell with two chain faces pairs them.
Passes are then followed from face to face into open chains (boundary to
boundary) and closed loops.
"""
from __future__ import annotations

from collections import deque
from typing import List, Sequence

import numpy as np

from .meshgen import BC_FREESTREAM, Mesh


class ChainError(ValueError):
    pass


def _cells(conn) -> List[List[int]]:
    """Corner lists (consecutive repeats dropped: a triangle may be stored as
    a quadrilateral with its last corner repeated)."""
    out = []
    for row in np.asarray(conn).tolist():
        c = [row[0]]
        for v in row[1:]:
            if v != c[-1] and v != c[0]:
                c.append(v)
        if len(c) not in (3, 4):
            raise ChainError(f"cell with {len(c)} corners")
        out.append(c)
    return out


def _ccw(x: np.ndarray, cells: List[List[int]]) -> List[List[int]]:
    out = []
    for c in cells:
        p = x[c]
        a = np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1])
        if a == 0:
            raise ChainError("degenerate cell")
        out.append(c if a > 0 else c[::-1])
    return out


def build_chains(x: np.ndarray, conn, loop_start_offset: int = 2, name: str = "unstructured") -> Mesh:
    """Faces, chains, colours and orphan faces of the mesh (x: (nNode, 2),
    conn: per cell 3 or 4 corner nodes, a triangle optionally with its last
    corner repeated).  Boundary faces come out as far-field (-1); the solver's
    preprocess tags those within radius 3 of the origin as walls
    (preprocess.cpp:176-186)."""
    x = np.asarray(x, np.float64)
    cells = _ccw(x, _cells(conn))
    ncell = len(cells)

    # ---- faces: edge (a, b) counter-clockwise around its first cell ----
    edge_of = {}
    f_nodes, f_cells = [], []          # per face: (n1, n2), [left, right]
    cell_faces = []                    # per cell: face ids in corner order
    for c, nodes in enumerate(cells):
        fl = []
        for k in range(len(nodes)):
            a, b = nodes[k], nodes[(k + 1) % len(nodes)]
            key = (min(a, b), max(a, b))
            f = edge_of.get(key)
            if f is None:
                f = len(f_nodes)
                edge_of[key] = f
                f_nodes.append((a, b))
                f_cells.append([c, BC_FREESTREAM])
            else:
                if f_cells[f][1] != BC_FREESTREAM:
                    raise ChainError("edge shared by more than two cells")
                if f_nodes[f] != (b, a):
                    raise ChainError("neighbouring cells with opposite orientation")
                f_cells[f][1] = c
            fl.append(f)
        cell_faces.append(fl)
    nface = len(f_nodes)
    interior = np.array([fc[1] >= 0 for fc in f_cells])

    def other(f, c):
        l, r = f_cells[f]
        return r if l == c else l

    # ---- virtual quadrilaterals: adjacent triangles paired, best-shaped union first ----
    tri = [len(cf) == 3 for cf in cell_faces]

    def union_sides(a, b, f):
        """sides of the union of triangles a and b (sharing face f),
        counter-clockwise: two of a's faces, then two of b's"""
        ka, kb = cell_faces[a].index(f), cell_faces[b].index(f)
        fa, fb = cell_faces[a], cell_faces[b]
        return [fa[(ka + 1) % 3], fa[(ka + 2) % 3], fb[(kb + 1) % 3], fb[(kb + 2) % 3]]

    def union_quality(a, b, f):
        """smallest interior angle's sine of the union (<= 0: not convex)"""
        sides = union_sides(a, b, f)
        pts = []
        for g, c in zip(sides, (a, a, b, b)):
            n1, n2 = f_nodes[g] if f_cells[g][0] == c else f_nodes[g][::-1]
            pts.append(n1)
        p = x[pts]
        q = 1.0
        for k in range(4):
            u = p[k] - p[k - 1]
            v = p[(k + 1) % 4] - p[k]
            cr = u[0] * v[1] - u[1] * v[0]
            q = min(q, cr / max(np.hypot(*u) * np.hypot(*v), 1e-300))
        return q

    cand = []
    for f in range(nface):
        l, r = f_cells[f]
        if r >= 0 and tri[l] and tri[r]:
            qv = union_quality(l, r, f)
            if qv > 0:
                cand.append((-qv, f))
    cand.sort()
    mate = [-1] * ncell
    mface = [-1] * ncell

    def join(l, r, f):
        mate[l], mate[r] = r, l
        mface[l] = mface[r] = f

    # 1. triangles whose longest edges coincide (the cut of a split quadrilateral)
    def longest(c):
        return max(cell_faces[c], key=lambda g: np.hypot(*(x[f_nodes[g][1]] - x[f_nodes[g][0]])))
    ok = {f for _, f in cand}
    for c in range(ncell):
        if tri[c] and mate[c] < 0:
            f = longest(c)
            n = other(f, c)
            if f in ok and n >= 0 and mate[n] < 0 and longest(n) == f:
                join(c, n, f)
    # 2. the rest greedily, best union first
    for _, f in cand:
        l, r = f_cells[f]
        if mate[l] < 0 and mate[r] < 0:
            join(l, r, f)
    # 3. augmenting alternating paths between triangles left unpaired
    def augment(u):
        seen = {u}
        stack = [(u, iter([g for g in cell_faces[u] if g in ok]))]
        path = []
        while stack:
            c, it = stack[-1]
            g = next(it, None)
            if g is None:
                stack.pop()
                if path:
                    path.pop()
                continue
            n = other(g, c)
            if n in seen:
                continue
            seen.add(n)
            if mate[n] < 0:
                path.append((c, n, g))
                for a, b, h in path:
                    join(a, b, h)
                return True
            m2 = mate[n]
            if m2 in seen:
                continue
            seen.add(m2)
            path.append((c, n, g))
            stack.append((m2, iter([h for h in cell_faces[m2] if h in ok and h != mface[m2]])))
        return False
    for c in range(ncell):
        if tri[c] and mate[c] < 0:
            augment(c)
    diag = {min(c, mate[c]): mface[c] for c in range(ncell) if mate[c] >= 0}
    # virtual cells: (members, sides); a real quad is its own, an unpaired
    # triangle has three sides and carries no chain
    vmembers, vsides, vof = [], [], [-1] * ncell
    for c in range(ncell):
        if vof[c] >= 0:
            continue
        if not tri[c]:
            members, sides = [c], list(cell_faces[c])
        elif mate[c] >= 0:
            m = mate[c]
            members, sides = [c, m], union_sides(c, m, diag[min(c, m)])
        else:
            members, sides = [c], list(cell_faces[c])
        for mcell in members:
            vof[mcell] = len(vmembers)
        vmembers.append(members)
        vsides.append(sides)
    nv = len(vmembers)

    def vother(f, v):
        """virtual cell across face f from v (-1: boundary)"""
        l, r = f_cells[f]
        c = r if vof[l] == v else l
        return vof[c] if c >= 0 else -1

    # ---- strips through opposite sides of (virtual) quadrilaterals ----
    strip_of = {}            # (virtual cell, pair 0/1) -> strip id
    strips = []              # list of [(v, f_in, f_out)], ends: (kind0, kind1) with 'b' boundary, 'd' dead end
    ends = []

    def pair_of(v, f):
        return vsides[v].index(f) % 2

    def trace(v, f_in):
        seq = []
        while True:
            sd = vsides[v]
            k = sd.index(f_in)
            f_out = sd[(k + 2) % 4]
            seq.append((v, f_in, f_out))
            strip_of[(v, k % 2)] = len(strips)
            if not interior[f_out]:
                return seq, "b"
            n = vother(f_out, v)
            if len(vsides[n]) == 3:
                return seq, "d"
            if (n, pair_of(n, f_out)) in strip_of:
                return seq, "c"     # back at the start: closed
            v, f_in = n, f_out

    def start_kind(v, f_in):
        if not interior[f_in]:
            return "b"
        n = vother(f_in, v)
        return "d" if len(vsides[n]) == 3 else "x"

    for f in range(nface):                   # strips that start at a boundary or a dead end
        for c in f_cells[f]:
            if c < 0:
                continue
            v = vof[c]
            if len(vsides[v]) != 4 or f not in vsides[v] or (v, pair_of(v, f)) in strip_of:
                continue
            k0 = start_kind(v, f)
            if k0 == "x":
                continue
            seq, k1 = trace(v, f)
            strips.append(seq)
            ends.append((k0, k1))
    for v in range(nv):                      # closed strips
        if len(vsides[v]) != 4:
            continue
        for k in (0, 1):
            if (v, k) not in strip_of:
                seq, k1 = trace(v, vsides[v][k])
                if k1 != "c":
                    raise ChainError("internal: strip without ends is not closed")
                strips.append(seq)
                ends.append(("c", "c"))
    ns = len(strips)

    # ---- keep or orphan each strip ----
    # forced: a boundary end keeps it; a dead end or a cell crossed twice orphans it
    forced = [-1] * ns
    for i, seq in enumerate(strips):
        keep = "b" in ends[i]
        drop = "d" in ends[i] or len({v for v, _, _ in seq}) != len(seq)
        if keep and drop:
            raise ChainError("a line from the boundary ends at an unpaired triangle or crosses itself")
        forced[i] = 1 if keep else (0 if drop else -1)
    # a virtual quadrilateral keeps exactly one of its two strips
    xor_adj = [[] for _ in range(ns)]
    for v in range(nv):
        if len(vmembers[v]) == 2:
            a, b = strip_of[(v, 0)], strip_of[(v, 1)]
            if a == b:
                raise ChainError("one line crosses a triangle pair twice")
            xor_adj[a].append(b)
            xor_adj[b].append(a)
    val = [-1] * ns
    for i in sorted(range(ns), key=lambda i: forced[i] < 0):
        if val[i] >= 0:
            continue
        comp, stack = [i], [i]
        tmp = {i: 0}
        while stack:
            a = stack.pop()
            for b in xor_adj[a]:
                if b not in tmp:
                    tmp[b] = 1 - tmp[a]
                    comp.append(b)
                    stack.append(b)
                elif tmp[b] == tmp[a]:
                    raise ChainError("triangle pairs force a line to be kept and dropped at once")
        # the two assignments of the component; take one that honours every forced strip,
        # the one keeping more cells on chains when both do
        best = None
        for flipv in (0, 1):
            assign = {a: tmp[a] ^ flipv for a in comp}
            if any(forced[a] >= 0 and forced[a] != assign[a] for a in comp):
                continue
            kept = sum(len(strips[a]) for a in comp if assign[a])
            if best is None or kept > best[0]:
                best = (kept, assign)
        if best is None:
            raise ChainError("triangle pairs force a line from the boundary to be dropped")
        for a, t in best[1].items():
            val[a] = t
    orphan = np.zeros(nface, bool)
    for i, seq in enumerate(strips):
        if not val[i]:
            for _, f_in, f_out in seq:
                if interior[f_in]:
                    orphan[f_in] = True
                if interior[f_out]:
                    orphan[f_out] = True
    for v in range(nv):
        if len(vsides[v]) == 3:
            for f in vsides[v]:
                if not interior[f]:
                    raise ChainError("an unpaired triangle on the boundary")
                orphan[f] = True

    # ---- chains: kept strips, through both triangles of a pair ----
    chains = []   # (cells, faces, closed); faces[k] enters cells[k]
    for i, seq in enumerate(strips):
        if not val[i]:
            continue
        cs, fs = [], [seq[0][1]]
        for v, f_in, f_out in seq:
            mem = vmembers[v]
            if len(mem) == 1:
                cs.append(mem[0])
            else:
                a = mem[0] if f_in in cell_faces[mem[0]] else mem[1]
                b = mem[1] if a == mem[0] else mem[0]
                cs += [a, b]
                fs.append(diag[min(a, b)])
            fs.append(f_out)
        closed = ends[i] == ("c", "c")
        if closed:
            fs = fs[:-1]
            if len(cs) < 3:
                raise ChainError("closed chain with fewer than 3 cells")
        chains.append((cs, fs, closed))

    # ---- colours: greedy, closed loops first, no two chains of one colour share a cell ----
    order = sorted(range(len(chains)), key=lambda i: (not chains[i][2], i))
    chains_of_cell = [[] for _ in range(ncell)]
    for i, (cs, _, _) in enumerate(chains):
        for c in cs:
            chains_of_cell[c].append(i)
    colour = [-1] * len(chains)
    for i in order:
        taken = {colour[j] for c in chains[i][0] for j in chains_of_cell[c] if colour[j] >= 0}
        k = 0
        while k in taken:
            k += 1
        colour[i] = k
    ncol = max(colour) + 1 if chains else 0
    order = sorted(range(len(chains)), key=lambda i: (colour[i], order.index(i)))

    # ---- cell 0: the first cell of an open chain whose other chain (if any)
    # is closed and long enough to hold it away from its ends ----
    pos = [dict() for _ in range(ncell)]
    for i, (cs, _, _) in enumerate(chains):
        for k, c in enumerate(cs):
            pos[c][i] = k
    pick = -1
    for i in order:
        cs, _, closed = chains[i]
        if closed or len(cs) < 7:    # cell 0 first, and at least 3 cells from the far end
            continue
        c = cs[0]
        if all(chains[j][2] and len(chains[j][0]) >= 2 * loop_start_offset + 2 for j in pos[c] if j != i):
            pick = c
            break
    if pick < 0:
        pick = chains[order[0]][0][len(chains[order[0]][0]) // 2] if chains else 0
    perm = np.arange(ncell)
    perm[0], perm[pick] = pick, 0           # new id of old cell
    # closed chains holding cell `pick` start loop_start_offset cells after it
    rot = {}
    for j in pos[pick]:
        if chains[j][2]:
            rot[j] = (pos[pick][j] - loop_start_offset) % len(chains[j][0])

    # ---- faces, oriented along their chains: face k has cell k on its left ----
    faces = np.empty((nface, 6), np.int64)
    for f in range(nface):
        (a, b), (l, r) = f_nodes[f], f_cells[f]
        faces[f] = (a, b, l, 0, r, 0 if r >= 0 else -1)

    def orient(f, left_cell):
        if faces[f, 2] != left_cell:
            a, b, l, el, r, er = faces[f]
            faces[f] = (b, a, r, er, l, el)

    chain_conn, lens = [], []
    for i in order:
        cs, fs, closed = chains[i]
        if closed:
            s = rot.get(i, 0)
            cs = cs[s:] + cs[:s]
            fs = fs[s:] + fs[:s]
            for k, f in enumerate(fs):
                orient(f, cs[k])
            ids = fs + [fs[0]]
        else:
            for k in range(1, len(cs)):
                orient(fs[k], cs[k])
            ids = fs
        chain_conn.extend(ids)
        lens.append(len(ids))
    start = np.zeros(len(lens) + 1, np.int32)
    start[1:] = np.cumsum(lens)
    per_colour = [sum(1 for i in order if colour[i] == k) for k in range(ncol)]

    # ---- renumber cell 0 ----
    cf = faces[:, [2, 4]]
    faces[:, 2] = np.where(cf[:, 0] >= 0, perm[np.maximum(cf[:, 0], 0)], cf[:, 0])
    faces[:, 4] = np.where(cf[:, 1] >= 0, perm[np.maximum(cf[:, 1], 0)], cf[:, 1])
    conn_out = np.empty((ncell, 4), np.int32)
    for c, nodes in enumerate(cells):
        conn_out[perm[c]] = nodes + [nodes[-1]] * (4 - len(nodes))
    orphans = np.nonzero(orphan)[0].astype(np.int32)
    return Mesh(nc=0, nr=0, x=x.copy(), conn=conn_out, faces=faces.astype(np.int32), chain_start=start,
                chain_conn=np.asarray(chain_conn, np.int32), chains_per_color=per_colour, name=name,
                orphans=orphans)


def validate(mesh: Mesh, order: int = 5) -> None:
    """Raises ChainError unless the mesh meets the layout above: faces
    counter-clockwise around their left cell, the chain contract, every face
    on exactly one chain or in the orphan list (orphans interior), colours
    cell-disjoint and listed in order, and cell 0 clear of the ghost quirk for
    nGhost = 3 (order 5) or 2 (computeRHS.cpp:150-175)."""
    x = mesh.x
    faces = mesh.faces
    ng = 3 if order == 5 else 2
    cnt = np.zeros(mesh.n_face, int)
    per_chain_cells = []
    for i in range(mesh.n_chain):
        ids = mesh.chain_conn[mesh.chain_start[i]:mesh.chain_start[i + 1]]
        closed = ids[0] == ids[-1]
        body = ids[:-1] if closed else ids
        np.add.at(cnt, body, 1)
        cells = [int(faces[f, 2]) for f in ids[:-1]]
        for k, f in enumerate(body):
            if k < len(cells) and faces[f, 2] != cells[k]:
                raise ChainError(f"chain {i}: face {k} has the wrong left cell")
            if closed and faces[f, 4] != cells[k - 1]:
                raise ChainError(f"chain {i}: face {k} has the wrong right cell")
            if not closed and 0 < k < len(cells) and faces[f, 4] != cells[k - 1]:
                raise ChainError(f"chain {i}: face {k} has the wrong right cell")
        if not closed:
            if faces[ids[0], 4] >= 0 or faces[ids[-1], 4] >= 0:
                raise ChainError(f"open chain {i} does not end on boundary faces")
            if faces[ids[-1], 2] != cells[-1]:
                raise ChainError(f"open chain {i}: last face's left cell")
            # the mirrored ghosts of the cells at positions 1..ng-1 from either
            # end would read as cell 0 (-0 == 0); the end cells are special-cased
            if 0 in cells:
                p = cells.index(0)
                if 1 <= p <= ng - 1 or 1 <= len(cells) - 1 - p <= ng - 1:
                    raise ChainError(f"cell 0 next to an end of open chain {i}")
        elif 0 in cells:
            # the periodic wrap copies at the ghost positions would read as ghosts
            p = cells.index(0)
            if p == 0 or p == len(cells) - 1:
                raise ChainError(f"cell 0 first or last in closed chain {i}")
        if len(set(cells)) != len(cells):
            raise ChainError(f"chain {i} visits a cell twice")
        per_chain_cells.append(set(cells))
    orph = np.asarray(mesh.orphans, int)
    np.add.at(cnt, orph, 1)
    if cnt.min() != 1 or cnt.max() != 1:
        raise ChainError("every face must lie on exactly one chain or be an orphan")
    if orph.size and (faces[orph, 4] < 0).any():
        raise ChainError("orphan faces must be interior")
    k0 = 0
    for k, n in enumerate(mesh.chains_per_color):
        seen = set()
        for i in range(k0, k0 + n):
            if seen & per_chain_cells[i]:
                raise ChainError(f"colour {k}: chains share a cell")
            seen |= per_chain_cells[i]
        k0 += n
    if k0 != mesh.n_chain:
        raise ChainError("colour counts do not add up to the chain count")
    for f, (n1, n2, cl, el, cr, er) in enumerate(faces):
        nodes = [v for v in dict.fromkeys(mesh.conn[cl].tolist())]
        k = nodes.index(n1) if n1 in nodes else -1
        if k < 0 or nodes[(k + 1) % len(nodes)] != n2:
            raise ChainError(f"face {f} is not counter-clockwise around its left cell")
