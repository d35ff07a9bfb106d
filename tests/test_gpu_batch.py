"""Concurrent multi-cell optimisation (dftpy_amd/batch.py): cells of uneven
grid sizes on several streams give the same results as one after another
(and with the device-driven, graph-launched TN as with the host-driven one),
every cell reaches the reference stop rule, and each FP32 result's FP64
energy lies within 1e-6 Ha/atom of the cell's FP64 ground state."""
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


def test_batch_graph_tn_matches_host_tn():
    """The batch with the device-driven TN (one captured hipGraph per plan,
    updated in place from cell to cell, captures in several threads) gives
    the host-driven TN's iteration counts cell by cell and its energies to
    1e-8 Ha/atom (the two differ in the last bits of cos/sin of the line
    search step, which FP32 fields can carry into the trajectory)."""
    from dftpy_amd.batch import make_cells, run_batch
    cells = make_cells(6, seed=4)
    res, _ = run_batch(cells, workers=3)
    resg, _ = run_batch(cells, workers=3, graph=True)
    print([(r["nr"], r["niter"], r["evals"]) for r in resg])
    for a, b in zip(res, resg):
        assert a["status"] == b["status"] == 0
        assert a["niter"] == b["niter"] and a["evals"] == b["evals"]
        assert abs(a["energy"] - b["energy"]) < 1e-8 * a["nat"]


def test_batch_fp32_energies_vs_fp64():
    """Each cell's FP32-bulk result (learned start, preconditioned TN) against
    the FP64 TN of the reference algorithm from the uniform density converged
    to 1e-9 Ha/atom: the FP64 energy of the FP32 density within 1e-6 Ha/atom."""
    from dftpy_amd.batch import fp64_reference, make_cells, run_batch
    cells = make_cells(6, seed=3)
    res, wall = run_batch(cells, workers=3, check64=True)
    for (ions, nr), r in zip(cells, res):
        e_ref, status = fp64_reference(ions, nr)
        err = (r["energy64"] - e_ref) / r["nat"]
        print("%3d atoms %s: FP32 status %d, %d evals; FP64 energy of the result %.10f, FP64 TN %.10f (status %d): "
              "%.2e Ha/atom" % (r["nat"], tuple(nr), r["status"], r["evals"], r["energy64"], e_ref, status, err),
              flush=True)
        assert status == 0 and r["status"] == 0
        assert abs(err) <= 1e-6
