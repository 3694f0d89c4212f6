"""The row-bundle key's y / ty by a 32-bit multiply-high (pk_scatter.hip make_tile_geom,
pk_particles.h row_tile_key): exact for every row-bundle height and row index of the
supported meshes (n <= 4096)."""
import numpy as np


def test_multiply_high_division_is_exact():
    hy = np.arange(4096, dtype=np.uint64)
    for ty in range(2, 4097):
        mul = np.uint64(((1 << 32) - 1 + ty) // ty)   # ceil(2^32 / ty), as make_tile_geom
        assert int(mul) < (1 << 32)
        q = (hy * mul) >> np.uint64(32)
        np.testing.assert_array_equal(q, hy // np.uint64(ty), err_msg=f"ty={ty}")
