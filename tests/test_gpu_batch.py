"""Batches of independent slabs of different sizes advanced together, one
hipGraph per iteration -- fused (each voxel-update pass of all slabs in one
launch per kernel variant) or one branch per slab: each slab's result is the
one its own xt_icd_iterate gives.  With one line in flight per frame every
slab's run is deterministic, so the comparison is exact."""
import dataclasses

import numpy as np
import pytest

from conftest import SMALL, TINY, make_scan
from timbir_python_amd import xt_hip

pytestmark = pytest.mark.gpu

SLABS = [dataclasses.replace(TINY, iterations=4),
         dataclasses.replace(SMALL, iterations=4),
         dataclasses.replace(TINY, name="tiny48", N_xy=48, N_p=64, seed=21, iterations=4)]


@pytest.mark.parametrize("fuse", [1, 0], ids=["fused", "branches"])
def test_batch_matches_single_slabs(scan_root, fuse):
    paths = [make_scan(c, scan_root, "batch")[0] for c in SLABS]
    saved = {o: xt_hip.get_option(o)
             for o in (xt_hip.OPT_LINES, xt_hip.OPT_RESYNC_EVERY, xt_hip.OPT_BATCH_FUSE)}
    xt_hip.set_option(xt_hip.OPT_BATCH_FUSE, fuse)
    xt_hip.set_option(xt_hip.OPT_LINES, 1)
    xt_hip.set_option(xt_hip.OPT_RESYNC_EVERY, 1)
    try:
        single = []
        for c, p in zip(SLABS, paths):
            with xt_hip.Engine(p, c.cli_args()) as eng:
                with xt_hip.Icd(eng) as icd:
                    icd.iterate(4)
                    icd.download()
                single.append(eng.volume().copy())
        engs = [xt_hip.Engine(p, c.cli_args()) for c, p in zip(SLABS, paths)]
        icds = [xt_hip.Icd(e) for e in engs]
        try:
            with xt_hip.IcdBatch(icds) as b:
                print(f"[gpu] batch of {len(icds)} slabs, fuse={fuse}", flush=True)
                b.iterate(1)
                b.iterate(3)
            batched = []
            for i, e in zip(icds, engs):
                assert i.stats().iterations == 4
                i.download()
                batched.append(e.volume().copy())
        finally:
            for i in icds:
                i.close()
            for e in engs:
                e.close()
    finally:
        for o, v in saved.items():
            xt_hip.set_option(o, v)
    for c, a, b in zip(SLABS, single, batched):
        print(f"  {c.name}: max |batch - single| = {np.abs(a - b).max():.3e}")
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)
