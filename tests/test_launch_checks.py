"""Host-side launch checks (CPU only): the grids the fused and step-wise
launches would use, and the geometries they refuse before anything reaches
the device."""
import pytest

from beep_amd import _lib


def test_fused_geometry(gpu_lib):
    assert list(_lib.check_launch(_lib.CHECK_FUSED, [4096, 512, 60000, 163840])) == [4096]
    for args, msg in [([0, 256, 100, 65536], "no jobs"),
                      ([10, 96, 100, 65536], "whole wavefronts"),      # not a multiple of 64
                      ([10, 1024, 100, 65536], "whole wavefronts"),    # above the kernels' bound of 512
                      ([10, 256, 70000, 65536], "dynamic LDS"),
                      ([2 ** 31, 64, 0, 65536], "dispatch limits")]:
        with pytest.raises(_lib.BeepError, match=msg) as e:
            _lib.check_launch(_lib.CHECK_FUSED, args)
        assert e.value.code == -1


def test_stepwise_geometry(gpu_lib):
    # 317 x 9 dense model, 1024 particles: Apad 320.
    g = _lib.check_launch(_lib.CHECK_STEPWISE, [317, 9, 320, 1024, 1024])
    cell_x, cell_y, dense_x, dense_y, sparse_x, sparse_y = g
    assert (cell_x, cell_y) == (1024, (317 * 9 + 255) // 256)
    # Dense tiles on a one-dimensional grid, a whole number of XCD rounds:
    # 144 row tiles x 5 column tiles = 720 = 8 x 90.
    assert (dense_x, dense_y) == (720, 1)
    g = _lib.check_launch(_lib.CHECK_STEPWISE, [317, 9, 320, 3, 3])     # 1 x 5 tiles: 8 workgroups, 3 empty
    assert (g[2], g[3]) == (8, 1)
    assert (sparse_x, sparse_y) == (1024 * 9, 2)
    # Jobs on x: more than 65535 jobs or rows are allowed.
    g = _lib.check_launch(_lib.CHECK_STEPWISE, [40, 9, 64, 100000, 100000])
    assert g[0] == 100000 and g[4] == 900000
    for args, msg in [([317, 9, 320, 1025, 1024], "more jobs than"),     # infectivity rows past the buffers
                      ([317, 9, 317, 10, 10], "multiple of 64"),
                      ([317, 9, 256, 10, 10], "multiple of 64"),         # padding below the area count
                      ([317, 9, 320, 0, 10], "no jobs"),
                      ([0, 9, 64, 1, 1], "empty model")]:
        with pytest.raises(_lib.BeepError, match=msg):
            _lib.check_launch(_lib.CHECK_STEPWISE, args)
    # A grid whose y dimension (blocks of cells) would pass 65535.
    with pytest.raises(_lib.BeepError, match="dispatch limits"):
        _lib.check_launch(_lib.CHECK_STEPWISE, [2 ** 22, 8, 2 ** 22, 1, 1])


def test_unknown_check(gpu_lib):
    with pytest.raises(_lib.BeepError, match="unknown check"):
        _lib.check_launch(7, [1, 2, 3])
