"""Concurrent multi-cell optimisation (dftpy_amd/batch.py): cells of uneven
grid sizes on several streams give the same results as one after another,
and every cell reaches the reference stop rule."""
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_sequential():
    from dftpy_amd.batch import make_cells, run_batch
    cells = make_cells(6, seed=3)
    assert len({nr for _, nr in cells}) >= 2  # uneven sizes
    res, wall = run_batch(cells, workers=3)
    res1, wall1 = run_batch(cells, workers=1)
    print(wall, wall1, [(r["nr"], r["niter"], r["evals"]) for r in res])
    for a, b in zip(res, res1):
        assert a["status"] == 0 and b["status"] == 0
        assert abs(a["energy"] - b["energy"]) < 1e-9 * a["nat"]
