"""Pins the CPU mesh checker (oracle/lenia_mesh_oracle.c, after mesh.cpp:34-2036) and the
library's default triangulation table (lh_mc_table) without a GPU.

The table is checked against an independent derivation written here from the cube's
topology alone and against the reference's EDGE_TABLE fixture (every crossing edge used,
no other).  The mesher is checked on what geometry fixes: closed, consistently oriented
surfaces with Euler characteristic 2 (sphere) and 0 (torus), outward normals, vertices
on the iso-surface of a linear field, a hand-counted octahedron, the per-field edge maps
of views 2-3 and the channel-shared edge map of the reference.
"""
import json
import os

import numpy as np
import pytest

from lenia_amd import engine as E
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
DX = [0, 1, 1, 0, 0, 1, 1, 0]
DY = [0, 0, 0, 0, 1, 1, 1, 1]
DZ = [0, 0, 1, 1, 0, 0, 1, 1]
EDGES = [(0, 1), (1, 2), (2, 3), (0, 3), (4, 5), (5, 6), (6, 7), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]


def derive_row(m):
    """Triangles of case m: each face's crossings joined (an ambiguous face separates
    its corners above iso), loops walked from their lowest edge, oriented so the fan
    normal points from the corners above iso to those below, fan-triangulated."""
    P = np.array([DX, DY, DZ]).T.astype(float)
    ins = lambda k: (m >> k) & 1
    edge_of = {frozenset(e): i for i, e in enumerate(EDGES)}
    adj = {}
    for ax in range(3):
        for side in (0, 1):
            cs = [k for k in range(8) if P[k][ax] == side]
            cyc = [cs[0]]
            rest = cs[1:]
            while rest:
                nxt = [k for k in rest if np.abs(P[k] - P[cyc[-1]]).sum() == 1][0]
                cyc.append(nxt)
                rest.remove(nxt)
            fe = [edge_of[frozenset((cyc[i], cyc[(i + 1) % 4]))] for i in range(4)]
            cr = [i for i in range(4) if ins(cyc[i]) != ins(cyc[(i + 1) % 4])]
            if len(cr) == 2:
                pairs = [(fe[cr[0]], fe[cr[1]])]
            elif len(cr) == 4:
                pairs = [(fe[3], fe[0]), (fe[1], fe[2])] if ins(cyc[0]) else [(fe[0], fe[1]), (fe[2], fe[3])]
            else:
                pairs = []
            for a, b in pairs:
                adj.setdefault(a, []).append(b)
                adj.setdefault(b, []).append(a)
    seen, out = set(), []
    for s in sorted(adj):
        if s in seen:
            continue
        loop, prev, cur = [s], None, s
        seen.add(s)
        while True:
            nb = [n for n in adj[cur] if n != prev and n not in seen]
            if not nb:
                break
            prev, cur = cur, nb[0]
            loop.append(cur)
            seen.add(cur)
        M = [(P[EDGES[e][0]] + P[EDGES[e][1]]) / 2 for e in loop]
        n = sum(np.cross(M[i] - M[0], M[i + 1] - M[0]) for i in range(1, len(loop) - 1))
        d = sum((P[b] - P[a]) if ins(a) else (P[a] - P[b]) for a, b in (EDGES[e] for e in loop))
        if n @ d < 0:
            loop = [loop[0]] + loop[1:][::-1]
        for i in range(1, len(loop) - 1):
            out += [loop[0], loop[i], loop[i + 1]]
    return out


@pytest.fixture(scope="module")
def table():
    return E.mc_table()


def test_table_matches_independent_derivation(table):
    for m in range(256):
        row = derive_row(m)
        assert list(table[m][:len(row)]) == row, m
        assert (table[m][len(row):] == -1).all(), m


def test_table_uses_exactly_the_reference_crossing_edges(table):
    ref = json.load(open(os.path.join(HERE, "golden", "edge_table.json")))["values"]
    for m in range(256):
        row = [e for e in table[m] if e >= 0]
        assert len(row) % 3 == 0 and len(row) <= 15
        used = 0
        for e in row:
            used |= 1 << int(e)
        assert used == ref[m], (m, hex(used), hex(ref[m]))
        # complementary cases cross the same edges
        assert ref[m] == ref[255 - m]


def surface_stats(v, idx):
    tris = idx.reshape(-1, 3)
    directed = {}
    for a, b, c in tris:
        for p, q in ((a, b), (b, c), (c, a)):
            directed[(p, q)] = directed.get((p, q), 0) + 1
    closed = all(cnt == 1 and directed.get((q, p), 0) == 1 for (p, q), cnt in directed.items())
    n_edges = len({tuple(sorted(k)) for k in directed})
    return len(v) - n_edges + len(tris), closed


def grid(n):
    return np.mgrid[0:n, 0:n, 0:n].astype(np.float64)  # z, y, x


def test_sphere_is_closed_with_outward_normals(table):
    z, y, x = grid(20)
    c = np.array([9.3, 10.1, 9.7])
    r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    f = np.clip(1 - r / 12, 0, 1).astype(np.float32)[None]
    v, idx = O.mesh([f], table, iso=0.5, cell=3.5)
    chi, closed = surface_stats(v, idx)
    assert closed and chi == 2
    pos = np.stack([v["x"], v["y"], v["z"]], 1) / 3.5
    nrm = np.stack([v["nx"], v["ny"], v["nz"]], 1)
    assert (((pos - c) * nrm).sum(1) > 0).all()
    # every vertex sits where the linear interpolation along its edge puts it: radius ~ 6
    assert np.abs(np.linalg.norm(pos - c, axis=1) - 6.0).max() < 0.35


def test_torus_has_euler_characteristic_zero(table):
    z, y, x = grid(24)
    rho = np.sqrt((x - 11.6) ** 2 + (y - 12.2) ** 2)
    d = np.sqrt((rho - 6.5) ** 2 + (z - 11.8) ** 2)
    f = np.clip(1 - d / 5, 0, 1).astype(np.float32)[None]
    v, idx = O.mesh([f], table, iso=0.5)
    chi, closed = surface_stats(v, idx)
    assert closed and chi == 0


def test_linear_field_vertices_on_the_plane(table):
    z, y, x = grid(12)
    f = (0.05 * x + 0.03 * y + 0.02 * z).astype(np.float32)[None]
    v, idx = O.mesh([f], table, iso=0.41, cell=1.0, face_only=True)
    assert len(v) > 0
    val = 0.05 * v["x"] + 0.03 * v["y"] + 0.02 * v["z"]
    inner = (v["x"] <= 10) & (v["y"] <= 10) & (v["z"] <= 10)  # away from the wrapped far faces
    assert np.abs(val[inner] - 0.41).max() < 1e-6


def test_single_cell_is_an_octahedron(table):
    f = np.zeros((1, 3, 3, 3), np.float32)
    f[0, 1, 1, 1] = 1.0
    v, idx = O.mesh([f], table, iso=0.5, cell=2.0)
    assert len(v) == 6 and len(idx) == 24
    pos = np.stack([v["x"], v["y"], v["z"]], 1)
    assert np.allclose(np.abs(pos - 2.0).sum(1), 1.0)  # half a cell from the centre, on the axes
    chi, closed = surface_stats(v, idx)
    assert closed and chi == 2
    assert (v["color"] == 0xFFFFFFFF).all()


def test_views_keep_one_edge_map_per_field(table):
    rs = np.random.default_rng(3)
    fs = [rs.random((2, 6, 7, 5), dtype=np.float32) for _ in range(3)]
    v3, i3 = O.mesh(fs, table, iso=0.5, colors=[0x11, 0x22])
    singles = [O.mesh([f], table, iso=0.5, colors=[0x11, 0x22]) for f in fs]
    assert len(v3) == sum(len(v) for v, _ in singles)
    assert len(i3) == sum(len(i) for _, i in singles)
    # the triangles of each field, as position triples, are those of that field alone
    pos3 = np.stack([v3["x"], v3["y"], v3["z"]], 1)
    tri3 = pos3[i3].reshape(-1, 9)
    allsingle = np.concatenate([np.stack([v["x"], v["y"], v["z"]], 1)[i].reshape(-1, 9) for v, i in singles])
    key = lambda a: sorted(map(tuple, a.tolist()))
    assert key(tri3) == key(allsingle)


def test_channels_share_the_edge_map(table):
    rs = np.random.default_rng(4)
    a = rs.random((1, 6, 6, 6), dtype=np.float32)
    f = np.concatenate([a, a])
    v, idx = O.mesh([f], table, iso=0.5, colors=[7, 9])
    v1, i1 = O.mesh([a], table, iso=0.5, colors=[7])
    assert len(v) == len(v1)  # channel 1 makes no vertex of its own
    assert (v["color"] == 7).all()
    assert np.array_equal(idx, np.concatenate([i1, i1]))


def test_reference_table_is_accepted_and_meshes_the_same_surface(table):
    """The reference's own TRIANGLE_TABLE (fixture written by make_tri_table.py): the
    same crossing edges as EDGE_TABLE, at most 5 triangles per case, a closed outward
    sphere, and the very vertices (positions) the default table gives."""
    ref_tab = np.array(json.load(open(os.path.join(HERE, "golden", "tri_table.json")))["rows"], dtype=np.int32)
    edges = json.load(open(os.path.join(HERE, "golden", "edge_table.json")))["values"]
    for m in range(256):
        row = [int(e) for e in ref_tab[m] if e >= 0]
        assert len(row) % 3 == 0 and len(row) <= 15 and (ref_tab[m][len(row):] == -1).all()
        assert sum(1 << e for e in set(row)) == edges[m], m
    z, y, x = grid(20)
    c = np.array([9.3, 10.1, 9.7])
    r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    f = np.clip(1 - r / 12, 0, 1).astype(np.float32)[None]
    v, idx = O.mesh([f], ref_tab)
    chi, closed = surface_stats(v, idx)
    assert closed and chi == 2
    pos = np.stack([v["x"], v["y"], v["z"]], 1)
    nrm = np.stack([v["nx"], v["ny"], v["nz"]], 1)
    assert ((pos / 3.5 - c) * nrm).sum(1).min() > 0
    v2, _ = O.mesh([f], table)
    key = lambda a: sorted(map(tuple, np.stack([a["x"], a["y"], a["z"]], 1).view(np.uint32).tolist()))
    assert key(v) == key(v2)


VERTEX_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "lenia_hip.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lh_vertex), offsetof(lh_vertex, x), offsetof(lh_vertex, y),
           offsetof(lh_vertex, z), offsetof(lh_vertex, nx), offsetof(lh_vertex, ny), offsetof(lh_vertex, nz),
           offsetof(lh_vertex, color));
    return 0;
}
"""


def test_vertex_record_layout_matches_the_c_header(tmp_path):
    """lh_vertex (include/lenia_hip.h), the Python record and the checker's lo_vertex
    share the reference Vertex's 28-byte layout: x y z nx ny nz colour."""
    import subprocess
    src = tmp_path / "probe.c"
    src.write_text(VERTEX_PROBE)
    exe = tmp_path / "probe"
    try:
        subprocess.run(["gcc", "-I", os.path.join(HERE, "..", "include"), str(src), "-o", str(exe)], check=True,
                       capture_output=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"no C compiler: {e}")
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [E.VERTEX_DTYPE.itemsize] + [E.VERTEX_DTYPE.fields[n][1] for n in ("x", "y", "z", "nx", "ny", "nz", "color")]
    assert got == want == [28, 0, 4, 8, 12, 16, 20, 24]
    assert O.VERTEX_DTYPE == E.VERTEX_DTYPE
