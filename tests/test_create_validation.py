"""Context creation refuses invalid meshes with an error, before any device work (runs without a GPU).

The reference stops on a clockwise or degenerate triangle (selem_2d.c:75-80, djac < 1e-6: "Improperly
numbered triangle"); here adhk_sw2_create returns NULL and adhk_last_error names the triangle.  Node
indices outside the mesh or the instance, unknown boundary kinds, zero-length segments and inconsistent
ensemble offsets are refused the same way, so that no kernel ever sees them.
"""
import numpy as np
import pytest

import adh_kraken_amd as A
from adh_kraken_amd import _lib as L


def two_triangles():
    """Unit square split along its diagonal, counter-clockwise, with a depth segment on x = 1."""
    x = np.array([0.0, 1.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0, 1.0])
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    seg = np.array([[1, 2]], np.int32)
    return A.Mesh(x=x, y=y, z=np.zeros(4), tri=tri, seg=seg, seg_kind=np.array([L.BC_DEPTH], np.int32),
                  seg_val=np.array([[1.0, 1.0]]), mann=np.full(2, 0.03))


def refused(mesh, match, **kw):
    with pytest.raises(L.AdhkError, match=match):
        A.SModelSuper(mesh, **kw)


def test_clockwise_triangle_is_refused():
    m = two_triangles()
    m.tri[1] = [0, 3, 2]
    refused(m, r"improperly numbered triangle 1 \(nodes 0 3 2\)")


@pytest.mark.parametrize("scale", [0.0, 1e-4])
def test_degenerate_and_tiny_triangles_are_refused(scale):
    m = two_triangles()
    m.x = m.x * scale  # areas 0, then 5e-9: both below 1e-6
    m.y = m.y * scale
    refused(m, "improperly numbered triangle 0")


def test_node_index_out_of_range_is_refused():
    m = two_triangles()
    m.tri[0, 2] = 4
    refused(m, "triangle node index out of its instance's range")
    m = two_triangles()
    m.seg[0, 1] = -1
    refused(m, "segment node index out of its instance's range")


def test_unknown_boundary_kind_and_zero_length_segment_are_refused():
    m = two_triangles()
    m.seg_kind[0] = 7
    refused(m, "unknown segment kind 7")
    m = two_triangles()
    m.seg[0] = [2, 2]
    refused(m, "degenerate segment 0")


def test_ensemble_offsets_are_checked():
    m = two_triangles()
    # offsets that do not end at the totals
    refused(m, "instance offsets", instances=([0, 3], [0, 2], [0, 1]))
    # an element of instance 0 that uses a node of instance 1
    refused(m, "triangle node index out of its instance's range", instances=([0, 3, 4], [0, 1, 2], [0, 1, 1]))
    # an empty instance
    refused(m, "instance 1 is empty", instances=([0, 4, 4], [0, 2, 2], [0, 1, 1]))


def test_valid_mesh_passes_the_checks():
    """The checks accept the valid mesh: creation then goes on to the device, which is absent here."""
    if L.load() is None:
        pytest.skip("library missing")
    try:
        sm = A.SModelSuper(two_triangles())
    except L.AdhkError as e:  # no GPU in this process: the first HIP call reports it
        msg = str(e)
        assert "improperly" not in msg and "out of its instance" not in msg and "segment" not in msg
    else:
        sm.close()
