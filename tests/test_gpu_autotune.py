"""GPU: the autotuner changes only time.  Every candidate it tries must end
in the same state (it raises otherwise); the tuned settings then reproduce
the default settings' trajectory bit for bit, single case and batch."""
import pytest

from cudadsmc_amd import autotune, cases
from cudadsmc_amd.batch import Batch
from cudadsmc_amd.run import Run

import dsmc_testlib as T

pytestmark = pytest.mark.gpu


def test_autotune_run_keeps_results(gpu_lib):
    p, body = T.small_cylinder(80, 40, 12, 64, seed=41)
    p.sample_start = 0
    r = Run(p, body)
    r.initialize()
    r.solver(12, graph=False)
    cus = autotune.device_cus()
    assert cus >= 1
    best, table = autotune.tune(autotune.RunTarget(r), cus, steps=3, graph_steps_measure=32, log=None)
    assert len(table) == 5 + 3 + 4 + 4 + 4
    assert best["cell_tile"] in autotune.TILES and best["graph_steps"] in autotune.GRAPH_STEPS
    r.solver(20, graph=True)
    ref = Run(p, body)
    ref.initialize()
    ref.solver(32, graph=False)
    T.assert_same_particles(r.download(), ref.download(), what="tuned vs default ")
    r.close()
    ref.close()


def test_autotune_batch_keeps_results(gpu_lib):
    sw = cases.sweep(n=6, seed=9, particles=(2e4, 8e4), sample_from=0)
    b = Batch([c["params"] for c in sw], [c["body"] for c in sw])
    tgt = autotune.BatchTarget(b, lambda s: s.initialize(), warm_steps=5)
    best, table = autotune.tune(tgt, autotune.device_cus(), steps=3, graph_steps_measure=32, log=None)
    assert best["cell_blocks"] >= 1
    b.close()
