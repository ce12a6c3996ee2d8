"""O-grid mesh and chain generator in the reference's on-disk formats.

The reference ships no meshes and no generator; it only reads
``./input/domainNNN/{coord,conn,ncolors,iqloops,qloops,qedges}.dat`` and
``input.hamstr`` (io.cpp:13-51, io.cpp:138-268, io.cpp:338-368).  This module
builds body-fitted O-grids (unit circle, or a NACA-00xx thickness-formula
section) and their line/loop chains so that every face lies in exactly one
chain, chains are grouped by colour (all closed loops first, then all open
lines), and cell 0 never sits where the reference's ``-0 == 0`` ghost test
misfires (computeRHS.cpp:158,170).

Face records follow the reference's memory layout after ``readMeshData``:
``[n1, n2, cL, eL, cR, eR]`` (io.cpp:349-355) with the interior cell on the
left and ``n1 -> n2`` running counter-clockwise around the left cell, so the
normal ``(y2-y1, -(x2-x1))`` (preprocess.cpp:86-87) points from left to right.

Cells are numbered ``c = i + nc*j`` (``i`` circumferential, counter-clockwise,
``j`` radial from the wall) and nodes ``n = i + nc*j`` for ``j = 0..nr``.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional

import numpy as np

# Boundary ids (utils.h:92-96)
BC_FREESTREAM = -1
BC_WALL_INVISC = -2


@dataclasses.dataclass
class Mesh:
    """Mesh + chain description, as MESHBLOCK holds it after readMeshData."""
    nc: int
    nr: int
    x: np.ndarray            # (nNode, 2) float64
    conn: np.ndarray         # (nCell, 4) int32 corner nodes (counter-clockwise)
    faces: np.ndarray        # (nFace, 6) int32  [n1, n2, cL, eL, cR, eR]
    chain_start: np.ndarray  # (nChain+1,) int32 offsets into chain_conn
    chain_conn: np.ndarray   # (nChainFace,) int32 face ids
    chains_per_color: List[int]
    name: str = "ogrid"
    # interior faces in no chain (meshes with triangles; empty for all-quad meshes)
    orphans: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))

    @property
    def n_cell(self) -> int:
        return int(self.conn.shape[0])

    @property
    def n_face(self) -> int:
        return int(self.faces.shape[0])

    @property
    def n_node(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_chain(self) -> int:
        return int(self.chain_start.shape[0] - 1)


def geometric_spacing(n: int, total: float, ratio: float) -> np.ndarray:
    """n+1 monotone stations from 0 to total with constant growth ratio."""
    if abs(ratio - 1.0) < 1e-14:
        d0 = total / n
    else:
        d0 = total * (ratio - 1.0) / (ratio ** n - 1.0)
    s = np.zeros(n + 1)
    for j in range(n):
        s[j + 1] = s[j] + d0 * ratio ** j
    s[-1] = total
    return s


def ratio_for_first_spacing(n: int, total: float, first: float) -> float:
    """Growth ratio r with first*(r^n-1)/(r-1) == total (bisection)."""
    if first * n >= total:
        return 1.0
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        tot = first * (mid ** n - 1.0) / (mid - 1.0)
        if tot > total:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def circle_nodes(nc: int, nr: int, r_outer: float = 20.0, ratio: float = 1.08) -> np.ndarray:
    """Node grid (nc, nr+1, 2) of an O-grid from the unit circle to r_outer."""
    r = 1.0 + geometric_spacing(nr, r_outer - 1.0, ratio)
    th = 2.0 * math.pi * np.arange(nc) / nc
    xy = np.empty((nc, nr + 1, 2))
    xy[:, :, 0] = np.cos(th)[:, None] * r[None, :]
    xy[:, :, 1] = np.sin(th)[:, None] * r[None, :]
    return xy


def naca00xx_thickness(xc: np.ndarray, t: float, closed_te: bool = False) -> np.ndarray:
    """Half thickness of a NACA 4-digit symmetric section, chord 1.

    The standard formula (x^4 coefficient -0.1015) leaves a base of 0.021 t
    chord at the trailing edge; closed_te=True uses -0.1036, which closes it to
    a sharp edge.
    """
    xc = np.clip(xc, 0.0, 1.0)
    c4 = -0.1036 if closed_te else -0.1015
    return 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2
                      + 0.2843 * xc ** 3 + c4 * xc ** 4)


def naca_nodes(nc: int, nr: int, t: float = 0.12, r_outer: float = 20.0,
               first: Optional[float] = None, closed_te: bool = False) -> np.ndarray:
    """Node grid (nc, nr+1, 2): NACA-00xx surface to a circle of radius r_outer.

    Surface stations use cosine clustering at both edges; node i sits at
    parameter phi_i = 2*pi*i/nc, x = 0.5+0.5*cos(phi) (trailing edge at
    phi = 0, counter-clockwise: upper surface first).  Radial rays blend the
    surface point with the far-field point at the same phi, with geometric
    stretching whose first spacing defaults to 0.5/nr**1.5 chord.
    """
    phi = 2.0 * math.pi * np.arange(nc) / nc
    xs = 0.5 + 0.5 * np.cos(phi)
    ys = np.sign(np.sin(phi)) * naca00xx_thickness(xs, t, closed_te)
    body = np.stack([xs, ys], axis=1)
    outer = np.stack([0.5 + r_outer * np.cos(phi), r_outer * np.sin(phi)], axis=1)
    if first is None:
        first = 0.5 / nr ** 1.5
    ratio = ratio_for_first_spacing(nr, 1.0, first / r_outer)
    s = geometric_spacing(nr, 1.0, ratio)
    xy = body[:, None, :] + (outer - body)[:, None, :] * s[None, :, None]
    return xy


def build_ogrid(xy: np.ndarray, loop_start: int = 2, name: str = "ogrid") -> Mesh:
    """Topology, faces and chains of an O-grid whose nodes are xy[i, j].

    Chains: nr closed loops (one per radial layer j; the first face is listed
    again at the end) followed by nc open lines (one per circumferential
    index i, wall face first, far-field face last).  Loops start at column
    ``loop_start`` so that cell 0 = (0, 0) is neither the first nor the last
    cell of its loop.
    """
    nc, nr1, _ = xy.shape
    nr = nr1 - 1
    if nc < 5 or nr < 3:
        raise ValueError("O-grid needs nc >= 5 and nr >= 3")
    x = xy.transpose(1, 0, 2).reshape(-1, 2).copy()   # node i + nc*j

    I, J = np.meshgrid(np.arange(nc), np.arange(nr), indexing="xy")   # (nr, nc)

    def node(i, j):
        return (i % nc) + nc * j

    def cell(i, j):
        return (i % nc) + nc * j

    ncell = nc * nr
    conn = np.empty((ncell, 4), np.int32)
    # counter-clockwise around the cell (see module docstring)
    conn[cell(I, J).ravel()] = np.stack([node(I + 1, J), node(I, J), node(I, J + 1), node(I + 1, J + 1)],
                                        axis=-1).reshape(-1, 4)

    n_loop_faces = nc * nr
    n_line_faces = nc * (nr + 1)
    faces = np.empty((n_loop_faces + n_line_faces, 6), np.int32)
    # local edge ids: 0 inner, 1 at column i+1, 2 outer, 3 at column i
    K = I  # loop face at node column k of layer j
    faces[:n_loop_faces] = np.stack([node(K, J), node(K, J + 1), cell(K, J), np.full_like(K, 3),
                                     cell(K - 1, J), np.full_like(K, 1)], axis=-1).reshape(-1, 6)
    ii = np.arange(nc)[:, None]
    jj = np.arange(nr + 1)[None, :]
    lf = np.empty((nc, nr + 1, 6), np.int64)
    lf[..., 0] = node(ii + 1, jj)
    lf[..., 1] = node(ii, jj)
    lf[..., 2] = cell(ii, jj)
    lf[..., 3] = 0
    lf[..., 4] = cell(ii, jj - 1)
    lf[..., 5] = 2
    lf[:, 0, 4] = BC_FREESTREAM
    lf[:, 0, 5] = -1
    # far-field face: left is the outermost cell, n1 -> n2 counter-clockwise
    lf[:, nr, 0] = node(np.arange(nc), nr)
    lf[:, nr, 1] = node(np.arange(nc) + 1, nr)
    lf[:, nr, 2] = cell(np.arange(nc), nr - 1)
    lf[:, nr, 3] = 2
    lf[:, nr, 4] = BC_FREESTREAM
    lf[:, nr, 5] = -1
    faces[n_loop_faces:] = lf.reshape(-1, 6)

    ks = (loop_start + np.arange(nc)) % nc
    loops = np.arange(nr)[:, None] * nc + ks[None, :]
    loops = np.concatenate([loops, loops[:, :1]], axis=1)               # first face again
    lines = n_loop_faces + np.arange(nc)[:, None] * (nr + 1) + np.arange(nr + 1)[None, :]
    conn_list = np.concatenate([loops.ravel(), lines.ravel()]).astype(np.int32)
    lens = [nc + 1] * nr + [nr + 1] * nc
    start = np.zeros(len(lens) + 1, np.int32)
    start[1:] = np.cumsum(lens)
    return Mesh(nc=nc, nr=nr, x=x, conn=conn, faces=faces, chain_start=start,
                chain_conn=conn_list, chains_per_color=[nr, nc], name=name)


def build_ogrid_tri(xy: np.ndarray, tri_layers: int, loop_start: int = 2, name: str = "ogrid_tri") -> Mesh:
    """An O-grid whose outer ``tri_layers`` layers are split into triangles.

    Quad (i, j) of an outer layer is cut along the diagonal from node (i, j)
    to node (i+1, j+1) into an inner triangle (on ring face (i, j)) and an
    outer one (on ring face (i, j+1)).  A chain enters a cell through one face
    and leaves through another, so a triangle leaves one face uncovered:
    radial lines run ring face -> inner triangle -> diagonal -> outer triangle
    -> ring face, closed loops exist for the quad layers only, and the radial
    faces of the triangle layers (nc per layer, each between the outer
    triangle of column k and the inner triangle of column k-1) are orphan
    faces, which the solver treats face by face.

    Cells: quads c = i + nc*j for j < nq = nr - tri_layers, then per triangle
    layer and column the inner and the outer triangle.  Faces: loop faces of
    the quad layers, then per line its ring faces and diagonals in chain
    order, then the orphans.  conn lists a triangle's three corners
    counter-clockwise with the last one repeated.
    """
    nc, nr1, _ = xy.shape
    nr = nr1 - 1
    nq = nr - tri_layers
    if tri_layers <= 0:
        return build_ogrid(xy, loop_start, name)
    if nc < 5 or nq < 1 or nr < 3:
        raise ValueError("needs nc >= 5, nr >= 3 and at least one quad layer")
    x = xy.transpose(1, 0, 2).reshape(-1, 2).copy()

    def node(i, j):
        return (np.asarray(i) % nc) + nc * np.asarray(j)

    def quad(i, j):
        return (np.asarray(i) % nc) + nc * np.asarray(j)

    def tri_in(i, j):
        return nc * nq + 2 * ((np.asarray(i) % nc) + nc * (np.asarray(j) - nq))

    def tri_out(i, j):
        return tri_in(i, j) + 1

    def above(i, j):   # cell on the outer side of ring face (i, j)
        return quad(i, j) if j < nq else tri_in(i, j)

    def below(i, j):   # cell on the inner side of ring face (i, j), j >= 1
        return quad(i, j - 1) if j - 1 < nq else tri_out(i, j - 1)

    ncell = nc * nq + 2 * nc * tri_layers
    conn = np.empty((ncell, 4), np.int32)
    I, J = np.meshgrid(np.arange(nc), np.arange(nq), indexing="xy")
    conn[quad(I, J).ravel()] = np.stack([node(I + 1, J), node(I, J), node(I, J + 1), node(I + 1, J + 1)],
                                        axis=-1).reshape(-1, 4)
    It, Jt = np.meshgrid(np.arange(nc), np.arange(nq, nr), indexing="xy")
    a, b, c, d = node(It, Jt), node(It + 1, Jt), node(It + 1, Jt + 1), node(It, Jt + 1)
    conn[tri_in(It, Jt).ravel()] = np.stack([a, c, b, b], axis=-1).reshape(-1, 4)
    conn[tri_out(It, Jt).ravel()] = np.stack([a, d, c, c], axis=-1).reshape(-1, 4)

    faces = []
    # closed loops of the quad layers: radial face at node column k of layer j
    K = I
    loop_faces = np.stack([node(K, J), node(K, J + 1), quad(K, J), np.full_like(K, 3), quad(K - 1, J),
                           np.full_like(K, 1)], axis=-1).reshape(-1, 6)
    faces.append(loop_faces)
    n_loop = loop_faces.shape[0]
    # lines: ring faces and diagonals in chain order
    per_line = nr + 1 + tri_layers
    lines = np.empty((nc, per_line, 6), np.int64)
    for i in range(nc):
        rows = [[node(i + 1, 0), node(i, 0), above(i, 0), 0, BC_FREESTREAM, -1]]
        for j in range(1, nr + 1):
            if j - 1 >= nq:   # diagonal of layer j-1: outer triangle left, inner right
                rows.append([node(i + 1, j), node(i, j - 1), tri_out(i, j - 1), 0, tri_in(i, j - 1), 0])
            if j < nr:
                rows.append([node(i + 1, j), node(i, j), above(i, j), 0, below(i, j), 2])
            else:
                rows.append([node(i, nr), node(i + 1, nr), below(i, nr), 2, BC_FREESTREAM, -1])
        lines[i] = np.array(rows, np.int64)
    faces.append(lines.reshape(-1, 6))
    # orphans: radial faces of the triangle layers
    Ko, Jo = np.meshgrid(np.arange(nc), np.arange(nq, nr), indexing="xy")
    orph = np.stack([node(Ko, Jo), node(Ko, Jo + 1), tri_out(Ko, Jo), np.full_like(Ko, 2), tri_in(Ko - 1, Jo),
                     np.full_like(Ko, 1)], axis=-1).reshape(-1, 6)
    faces.append(orph)
    faces = np.concatenate(faces).astype(np.int32)
    n_line = nc * per_line

    ks = (loop_start + np.arange(nc)) % nc
    loops = np.arange(nq)[:, None] * nc + ks[None, :]
    loops = np.concatenate([loops, loops[:, :1]], axis=1)
    line_ids = n_loop + np.arange(n_line).reshape(nc, per_line)
    conn_list = np.concatenate([loops.ravel(), line_ids.ravel()]).astype(np.int32)
    lens = [nc + 1] * nq + [per_line] * nc
    start = np.zeros(len(lens) + 1, np.int32)
    start[1:] = np.cumsum(lens)
    orphans = (n_loop + n_line + np.arange(orph.shape[0])).astype(np.int32)
    return Mesh(nc=nc, nr=nr, x=x, conn=conn, faces=faces, chain_start=start, chain_conn=conn_list,
                chains_per_color=[nq, nc], name=name, orphans=orphans)


def circle_ogrid(nc: int = 128, nr: int = 64, r_outer: float = 20.0, ratio: float = 1.08,
                 tri_layers: int = 0) -> Mesh:
    name = f"circle{nc}x{nr}" + (f"_tri{tri_layers}" if tri_layers else "")
    return build_ogrid_tri(circle_nodes(nc, nr, r_outer, ratio), tri_layers, name=name)


def naca_ogrid(nc: int = 512, nr: int = 256, t: float = 0.12, r_outer: float = 20.0,
               first: Optional[float] = None, closed_te: bool = False, tri_layers: int = 0) -> Mesh:
    """NACA-00xx O-grid; tri_layers > 0 splits that many outer layers into
    triangles (build_ogrid_tri)."""
    name = f"naca{int(round(t*100)):02d}_{nc}x{nr}" + (f"_tri{tri_layers}" if tri_layers else "")
    return build_ogrid_tri(naca_nodes(nc, nr, t, r_outer, first, closed_te), tri_layers, name=name)


# ----------------------------------------------------------------------------
# reference file formats
# ----------------------------------------------------------------------------

def write_input_deck(path: str, mach: float, alpha: float, order: int, nsteps: int,
                     cfl: float, msweep: int = 1, scheme: str = "ADI", nwrite: int = 0) -> None:
    """``input.hamstr``: the 16 positional key=value lines of io.cpp:26-41."""
    lines = [
        f"Mach={float(mach)!r}", f"alpha={float(alpha)!r}", "beta=0.0", "rey=1000000.0",
        f"scheme={scheme}", "time integration=steady", f"order={order}",
        "timeacc=0", f"nsteps={nsteps}", f"nwrite={nwrite}", "dt=0.0", f"CFL={float(cfl)!r}",
        f"msweep={msweep}", "visc=0", "testcase=0", "output=0",
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_reference_case(mesh: Mesh, root: str, mach: float, alpha: float, order: int,
                         nsteps: int, cfl: float, msweep: int = 1, scheme: str = "ADI", nwrite: int = 0) -> str:
    """Write ``root/input.hamstr`` and ``root/input/domain000/*.dat``."""
    d = os.path.join(root, "input", "domain000")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "coord.dat"), "w") as fh:
        fh.write(f"{mesh.n_node}\n")
        for px, py in mesh.x:
            fh.write(f"{float(px)!r} {float(py)!r} 0.0\n")
    with open(os.path.join(d, "conn.dat"), "w") as fh:
        fh.write(f"{mesh.n_cell}\n")
        for row in mesh.conn:
            fh.write(" ".join(str(int(v)) for v in row) + "\n")
    with open(os.path.join(d, "ncolors.dat"), "w") as fh:
        fh.write(f"{len(mesh.chains_per_color)}\n")
        for v in mesh.chains_per_color:
            fh.write(f"{v}\n")
    with open(os.path.join(d, "iqloops.dat"), "w") as fh:
        fh.write(f"{mesh.chain_start.shape[0]}\n")
        for v in mesh.chain_start:
            fh.write(f"{int(v)}\n")
    with open(os.path.join(d, "qloops.dat"), "w") as fh:
        fh.write(f"{mesh.chain_conn.shape[0]}\n")
        fh.write("\n".join(str(int(v)) for v in mesh.chain_conn) + "\n")
    with open(os.path.join(d, "qedges.dat"), "w") as fh:
        fh.write(f"{mesh.n_face}\n")
        for n1, n2, c1, e1, c2, e2 in mesh.faces:
            # file order is n1 n2 c1 c2 e1 e2 (io.cpp:349-355)
            fh.write(f"{n1} {n2} {c1} {c2} {e1} {e2}\n")
    if mesh.orphans.size:
        # no counterpart in the reference: faces outside every chain
        with open(os.path.join(d, "orphans.dat"), "w") as fh:
            fh.write(f"{mesh.orphans.size}\n")
            fh.write("\n".join(str(int(v)) for v in mesh.orphans) + "\n")
    write_input_deck(os.path.join(root, "input.hamstr"), mach, alpha, order, nsteps, cfl, msweep, scheme, nwrite)
    return root


def read_reference_case(root: str) -> Mesh:
    """Parse the files written above with io.cpp's semantics.

    Reproduces the consecutive-duplicate removal of the chain offsets
    (io.cpp:229-245) and the left/right swap of faces whose first cell is
    -1 (io.cpp:360-365).
    """
    d = os.path.join(root, "input", "domain000")

    def ints(name):
        with open(os.path.join(d, name)) as fh:
            return [int(v) for v in fh.read().split()]

    with open(os.path.join(d, "coord.dat")) as fh:
        tok = fh.read().split()
    nn = int(tok[0])
    x = np.array([float(v) for v in tok[1:1 + 3 * nn]]).reshape(nn, 3)[:, :2].copy()
    c = ints("conn.dat")
    conn = np.array(c[1:1 + 4 * c[0]], np.int32).reshape(c[0], 4)
    col = ints("ncolors.dat")
    iq = ints("iqloops.dat")
    offs = []
    for v in iq[1:1 + iq[0]]:
        if offs and offs[-1] == v:
            continue
        offs.append(v)
    ql = ints("qloops.dat")
    chain_conn = np.array(ql[1:1 + ql[0]], np.int32)
    fe = ints("qedges.dat")
    nf = fe[0]
    raw = np.array(fe[1:1 + 6 * nf], np.int32).reshape(nf, 6)
    faces = np.empty_like(raw)
    faces[:, 0], faces[:, 1] = raw[:, 0], raw[:, 1]
    faces[:, 2], faces[:, 4] = raw[:, 2], raw[:, 3]
    faces[:, 3], faces[:, 5] = raw[:, 4], raw[:, 5]
    sw = faces[:, 2] == BC_FREESTREAM
    faces[sw] = faces[sw][:, [1, 0, 4, 5, 2, 3]]
    orph = np.zeros(0, np.int32)
    if os.path.exists(os.path.join(d, "orphans.dat")):
        ov = ints("orphans.dat")
        orph = np.array(ov[1:1 + ov[0]], np.int32)
    return Mesh(nc=0, nr=0, x=x, conn=conn, faces=faces,
                chain_start=np.array(offs, np.int32), chain_conn=chain_conn,
                chains_per_color=col[1:1 + col[0]], name=os.path.basename(root.rstrip("/")), orphans=orph)


def freestream(mach: float, alpha_deg: float, gamma: float = 1.4):
    """(rinf, uinf, vinf, einf) exactly as initFlow.cpp:107-142 computes them."""
    pi = 3.1415926535897932e+00
    deg2rad = pi / 180.0
    rinf = 1.0
    pinf = 1.0 / gamma
    uinf = mach * math.cos(alpha_deg * deg2rad)
    vinf = mach * math.sin(alpha_deg * deg2rad)
    gm1 = gamma - 1.0
    einf = pinf / gm1 + 0.5 * rinf * (uinf * uinf + vinf * vinf)
    return rinf, uinf, vinf, einf
