"""The cell filter's work lists built on the device (what the graph build
runs) against the host builder, which tests/test_lib_cpu.py pins on the CPU
against the definition of the needed tiles: the same items, tile lists and
covered pairs, array for array, for every work form, on random layouts with
empty cells, sparse to complete adjacency and both position orders."""
import numpy as np
import pytest

from conftest import flush_print
from test_lib_cpu import cell_work, random_cells

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,bt,seed,density", [(7, 8, 0, 0.15), (130, 8, 1, 0.15), (300, 24, 2, 0.15),
                                               (64, 16, 3, 0.5), (1500, 8, 4, 0.02), (600, 16, 5, 1.0),
                                               (2100, 12, 6, 0.05)])
def test_device_work_lists_equal_the_host_builder(m, bt, seed, density):
    from rnng_amd import _lib
    L = _lib.lib()
    cnt, adj, rng = random_cells(m, seed, density)
    for perm in (np.arange(m, dtype=np.int32), rng.permutation(m).astype(np.int32)):
        ntiles = int(((cnt + 31) // 32).sum())
        nrb = (ntiles + bt - 1) // bt
        for sched in (0, 1, 2):
            rbc = (rng.random(nrb) * 1000).astype(np.float32) if sched == 2 else None
            flush_print(f"m={m} bt={bt} density={density} sched={sched} identity={perm[0] == 0 and perm[-1] == m - 1}")
            sh, ih, ch, cov_h = cell_work(L, cnt, adj, perm, bt, sched, rbc)
            sd, idv, cd, cov_d = cell_work(L, cnt, adj, perm, bt, sched, rbc, device=True)
            assert np.array_equal(sh, sd)
            assert cov_h == cov_d
            assert np.array_equal(ih, idv), (ih[:5], idv[:5])
            assert np.array_equal(ch, cd)
