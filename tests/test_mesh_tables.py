"""Arena mesh tables: the host build matches the reference build's tables bit for bit, and the invariant the stackless
walk relies on holds (a parent's quantized box contains its subtree)."""
import numpy as np
import pytest

from conftest import oracle_or_skip, twin


@pytest.mark.parametrize("scale,fillet", [(1.0, 256.0), (0.8, 192.0)])
def test_mesh_tables_match_reference(scale, fillet):
    T = twin()
    R = oracle_or_skip()
    ref = R.RefSim(4, 2, scale, fillet)
    h = T.rgt_create(4, scale, fillet)
    try:
        tables = []
        for lib, handle in ((R.lib().rgref_mesh_tables, ref.h), (T.rgt_mesh_tables, h)):
            counts = np.zeros(3, np.int64)
            lib(handle, counts.ctypes.data, None, None, None, None)
            tris = np.zeros((counts[0], 16), np.float32)
            nodes = np.zeros((counts[1], 4), np.int32)
            meshes = np.zeros((counts[2], 15), np.float32)
            ints = np.zeros((counts[2], 4), np.int32)
            lib(handle, counts.ctypes.data, tris.ctypes.data, nodes.ctypes.data, meshes.ctypes.data, ints.ctypes.data)
            tables.append((counts, tris, nodes, meshes, ints))
        (c0, t0, n0, m0, i0), (c1, t1, n1, m1, i1) = tables
        assert np.array_equal(c0, c1)
        if scale == 1.0:
            assert c0[0] == 20332
        assert np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
        assert np.array_equal(n0, n1)
        assert np.array_equal(m0.view(np.uint32), m1.view(np.uint32))
        assert np.array_equal(i0, i1)
    finally:
        T.rgt_destroy(h)


@pytest.mark.parametrize("scale,fillet", [(1.0, 256.0), (0.8, 192.0)])
def test_bvh_parent_boxes_contain_subtrees(scale, fillet):
    """The walk skips a subtree when its root's box misses the query; that only equals testing every leaf when each
    parent box contains its subtree."""
    T = twin()
    h = T.rgt_create(1, scale, fillet)
    try:
        assert T.rgt_bvh_containment_violations(h) == 0
    finally:
        T.rgt_destroy(h)
