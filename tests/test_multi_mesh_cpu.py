"""The merged arrays of a mesh of several geometries (lp_set_geometries),
checked on the CPU through lp_debug_multi_mesh (no device):

* each geometry's threaded BVH, walked from its root (hit = first child,
  miss = next subtree), reaches exactly its own triangles, each once, with
  their original ids; the walks of all geometries cover every node once;
* the LDS-staged prefix holds the top of every tree: geometry g's root is
  the sum of the earlier geometries' top sizes, and its top nodes follow;
* each geometry's near-grid block (its cells and a sentinel, at its cell
  offset) lists only that geometry's triangles, and the blocks' list ranges
  follow one another;
* with the near grid off, the BVH is the same and no cells are returned.
"""
import ctypes as C

import numpy as np
import pytest

from lightpropagation_amd import _capi
from lightpropagation_amd import scenes as S

_U = C.POINTER(C.c_uint32)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)
_F = C.POINTER(C.c_float)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _build(scs, lds, r0, cell):
    merged, counts = S.geometries(scs)
    tri = np.ascontiguousarray(merged["tri"], np.float32).reshape(-1)
    med = np.ascontiguousarray(merged["med"], np.int32).reshape(-1)
    cnt = np.array(counts, np.int64)
    f = _capi.lib().lp_debug_multi_mesh
    f.restype = C.c_int
    f.argtypes = [_F, _I, _L, C.c_int32, C.c_int32, C.c_float, C.c_float, _L, _U, _I, _I, _U, _U, _I]
    sizes = np.zeros(6, np.int64)
    assert f(_ptr(tri, _F), _ptr(med, _I), _ptr(cnt, _L), len(cnt), lds, r0, cell, _ptr(sizes, _L),
             None, None, None, None, None, None) == 0
    nodes = np.zeros(8 * sizes[0], np.uint32)
    ids = np.zeros(sizes[1], np.int32)
    roots = np.zeros(len(cnt), np.int32)
    coffs = np.zeros(len(cnt), np.uint32)
    cells = np.zeros(max(2 * sizes[2], 1), np.uint32)
    lst = np.zeros(max(sizes[3], 1), np.int32)
    assert f(_ptr(tri, _F), _ptr(med, _I), _ptr(cnt, _L), len(cnt), lds, r0, cell, _ptr(sizes, _L),
             _ptr(nodes, _U), _ptr(ids, _I), _ptr(roots, _I), _ptr(coffs, _U), _ptr(cells, _U), _ptr(lst, _I)) == 0
    first = np.concatenate([[0], np.cumsum(counts)])
    return dict(sizes=sizes, nodes=nodes.reshape(-1, 8), ids=ids, roots=roots, coffs=coffs,
                cells=cells[:2 * sizes[2]].reshape(-1, 2), list=lst[:sizes[3]], first=first)


def _walk(nodes, root):
    hit = nodes[:, 3].view(np.int32)
    miss = nodes[:, 7].view(np.int32)
    n, visited, tris = int(root), [], []
    while n >= 0:
        assert len(visited) < len(nodes), "walk does not end"
        visited.append(n)
        h = int(hit[n])
        if h >= 0:
            n = h
        else:
            lf, cnt = h & 0xffffff, (h >> 24) & 0x7f
            tris.extend(range(lf, lf + cnt))
            n = int(miss[n])
    return visited, tris


def _scenes():
    return [S.maternal_fetal(fetal_depth=d, freq_maternal=f, freq_fetal=f, freq_torso=3)
            for d, f in ((2.5, 4), (3.5, 6), (3.0, 5))]


@pytest.mark.parametrize("lds", [64, 1000])
def test_merged_bvh_and_near_grid(lds):
    m = _build(_scenes(), lds, 0.5, 1.0)
    first, nodes = m["first"], m["nodes"]
    ng = len(m["roots"])
    seen = []
    top_base = 0
    for g in range(ng):
        visited, tris = _walk(nodes, m["roots"][g])
        own = list(range(first[g], first[g + 1]))
        assert sorted(tris) == own, g  # its own leaf-order triangles, each once
        assert sorted(m["ids"][tris].tolist()) == own, g  # with their original ids
        top = min(len(visited), lds)
        assert m["roots"][g] == top_base, g
        assert set(range(top_base, top_base + top)) <= set(visited), g  # its top follows the earlier tops
        top_base += top
        seen.extend(visited)
    assert sorted(seen) == list(range(len(nodes)))  # every node in exactly one tree
    # near-grid blocks: geometry after geometry, each with its sentinel
    per = int(m["sizes"][4])
    cells, lst = m["cells"], m["list"]
    assert per > 1 and len(cells) == ng * per
    prev_end = 0
    for g in range(ng):
        c0 = int(m["coffs"][g])
        assert c0 == g * per
        block = cells[c0:c0 + per, 0].astype(np.int64)
        assert block[0] == prev_end and np.all(np.diff(block) >= 0)
        entries = lst[block[0]:block[-1]]
        assert entries.size > 0 and entries.min() >= first[g] and entries.max() < first[g + 1], g
        prev_end = block[-1]
    assert prev_end == len(lst)


def test_without_near_grid():
    a = _build(_scenes(), 64, 0.5, 1.0)
    b = _build(_scenes(), 64, 0.0, 1.0)
    assert np.array_equal(a["nodes"], b["nodes"]) and np.array_equal(a["ids"], b["ids"])
    assert np.array_equal(a["roots"], b["roots"])
    assert b["sizes"][2] == 0 and b["sizes"][3] == 0 and np.all(b["coffs"] == 0)


def test_shared_frame_fits_every_geometry():
    # the shared near-grid frame is fitted to the vertices each geometry's
    # grid sees (triangle records in float); fitted to the input vertices, it
    # missed one geometry's box by a rounding at this radius
    GEOMS = [(1.5, 8, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, 10, {}), (3.5, 6, {})]
    scs = [S.maternal_fetal(fetal_depth=d, freq_maternal=f, freq_fetal=f, freq_torso=4, **kw) for d, f, kw in GEOMS]
    for r0 in (0.05, 0.07):
        r = _build(scs, 85, r0, 2 * r0)
        assert r["sizes"][4] > 0
