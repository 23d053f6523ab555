"""The threaded float32 CPU path the bench times as its baseline
(oracle/cpu_baseline.py) is the oracle's float32 solver with MKL transforms
and exposure threads: against RefSolver(f32) it must give the same FreezeMu
rows and an estimate within float32 transform rounding, for one worker and
for several (the worker-order sum of partial spectra, updc_core.c:1280-1285)."""
import numpy as np
import pytest

from conftest import oracle_exposures


@pytest.mark.parametrize("workers,threads", [(1, 1), (2, 4), (3, 2)])
def test_threaded_path_matches_the_oracle(small_ds, workers, threads):
    from oracle import cpu_baseline as B
    from oracle import oracle as O
    ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f32")
    rows_o, _ = ref.iterate(3)
    path_ref = O.RefSolver(small_ds.W, small_ds.H, oracle_exposures(small_ds), "f32")
    path = B.ThreadedF32(path_ref, workers, threads)
    rows = path.iterate(3)
    path.close()
    assert [r[:4] for r in rows] == [r[:4] for r in rows_o]
    st = O.compare_stats(ref.x, path_ref.x)
    assert st["norm_linf"] < 2e-5, st


def test_host_description():
    from oracle import cpu_baseline as B
    assert 1 <= B.host_cores() <= 16
    assert isinstance(B.cpu_model(), str)
