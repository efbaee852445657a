"""Snapshots and restart files on the host (no GPU): the reference's
snapshot! and restart_from_height examples (src/measures.jl:65-89,
src/initialvalues.jl:337-355), with .npz standing in for JLD2."""
import numpy as np

import swalbe as S


def test_snapshot_rows_follow_dumping():
    # src/measures.jl:79-88: h1 = reshape(1:25, 5, 5), h2 = reshape(5:5:125, 5, 5)
    h1 = np.arange(1, 26, dtype=float).reshape((5, 5), order="F")
    h2 = np.arange(5, 126, 5, dtype=float).reshape((5, 5), order="F")
    snap = np.zeros((2, 25))
    S.snapshot(snap, h1, 10, dumping=10)
    S.snapshot(snap, h2, 20, dumping=10)
    assert np.all(h1 == snap[0].reshape((5, 5), order="F"))
    assert np.all(h2 == snap[1].reshape((5, 5), order="F"))
    S.snapshot(snap, h1, 15, dumping=10)  # not a multiple of dumping: no row changes
    assert np.all(h2 == snap[1].reshape((5, 5), order="F"))


def test_restart_from_height_reads_the_requested_step(tmp_path):
    # src/initialvalues.jl:337-355: f["h_1"] = vec(h1); f["h_2"] = vec(h2)
    rng = np.random.default_rng(3)
    h1, h2 = rng.random((10, 10)), rng.random((10, 10))
    path = str(tmp_path / "file.npz")
    S.save_snapshots(path, {"h_1": h1, "h_2": h2})
    assert np.all(S.restart_from_height(path, timestep=1, size=(10, 10)) == h1)
    assert np.all(S.restart_from_height(path, timestep=2, size=(10, 10)) == h2)
    # timestep 0: the last column of the file
    assert np.all(S.restart_from_height(path, size=(10, 10)) == h2)
    # non-square, column-major layout
    h3 = rng.random((7, 4))
    S.save_snapshots(path, {"h_100": h3, "h_25": h1})
    assert np.all(S.restart_from_height(path, size=(7, 4)) == h3)
