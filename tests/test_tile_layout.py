"""Host checks of the tile-kernel eligibility test (no GPU needed).

The fused tile kernel hard-codes the structured face pattern of a Cartesian
mesh; eligibility is decided on the host from the mesh tables.  Every mesh
shape and boundary combination the tile kernel's GPU tests use must qualify,
and meshes whose element counts do not divide into tiles must not.
"""
import pytest

from dg_amd import FARFIELD, MULTIFLUID, PERIODIC, REFLECTIVE, SINGLEFLUID, Mesh

BCS_1D = [(FARFIELD, REFLECTIVE), (PERIODIC, PERIODIC), (REFLECTIVE, FARFIELD)]
BCS_2D = [(PERIODIC, PERIODIC, REFLECTIVE, FARFIELD), (FARFIELD, REFLECTIVE, PERIODIC, PERIODIC),
          (REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE), (PERIODIC, PERIODIC, PERIODIC, PERIODIC)]


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("ph", [SINGLEFLUID, MULTIFLUID])
def test_cartesian_meshes_qualify(order, ph):
    for bc in BCS_1D:
        for n in (128, 256, 1024):
            assert Mesh(1, order, n, 1, -1.0, 1.0, bc=bc, nfields=3 + ph).tileable, (bc, n)
    for bc in BCS_2D:
        for n in ((8, 4), (16, 8), (32, 16), (24, 12)):
            assert Mesh(2, order, n[0], n[1], 0.0, 1.6, 0.0, 0.6, bc=bc, nfields=4 + ph).tileable, (bc, n)


def test_ragged_meshes_do_not_qualify():
    assert not Mesh(1, 2, 203, 1, -1.0, 1.0, bc=BCS_1D[0], nfields=4).tileable
    assert Mesh(1, 2, 200, 1, -1.0, 1.0, bc=BCS_1D[0], nfields=4).tileable  # tiles of 8 elements
    assert not Mesh(2, 2, 12, 7, 0.0, 1.0, 0.0, 1.0, bc=BCS_2D[0], nfields=5).tileable
    assert not Mesh(2, 2, 16, 6, 0.0, 1.0, 0.0, 1.0, bc=BCS_2D[0], nfields=5).tileable
