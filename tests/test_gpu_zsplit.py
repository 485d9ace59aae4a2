"""The z transforms split by parity (codesthesis_dir_amd/csrc/vv_zsplit.h) in
the three 3D solvers: the tank, the annulus of rot3dmg.f and the barrier
sector.  By default the split runs from 96 z rows; VV_ZSPLIT=1 forces it at
every size, so that the solve tests of the three solvers (against the
oracles' dense or exact solves, and the reference's residual operators on
grids beyond them) run on it:

* small grids with odd and even z row counts (the middle row of odd counts
  belongs to the even half alone) on the usual 64 x 64 tiles;
* z row counts above 128, where the halves take the 64 x 128 and 128 x 64
  tiles, through the residual operators.
"""
import pytest

import test_gpu_3d as T3
import test_gpu_sector as TS
import test_gpu_tank as TT

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _split(monkeypatch):
    monkeypatch.setenv("VV_ZSPLIT", "1")


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(8, 6, 10), (7, 9, 6)])
def test_tank_poisson_split_matches_dense_solve(which, shape):
    TT.test_poisson_matches_dense_solve(which, shape)


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(8, 10, 6), (7, 6, 8)])
def test_rot3d_poisson_split_matches_exact_solve(which, shape):
    T3.test_poisson_matches_exact_solve(which, shape)


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(33, 65, 10), (12, 131, 8), (10, 150, 16)])
def test_rot3d_poisson_split_residual(which, shape):
    T3.test_poisson_residual_at_larger_sizes(which, shape)


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(6, 5, 8), (9, 7, 12)])
def test_sector_poisson_split_matches_dense_solve(which, shape):
    TS.test_poisson_matches_dense_solve(which, shape)


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(6, 131, 8), (5, 150, 9)])
def test_sector_poisson_split_residual(which, shape):
    TS.test_poisson_residual_at_size_limits(which, shape)
