"""Real-data ingest on the GPU (xt_counts_to_sinogram) against the C
restatement of gen_projection_4m_HDF's arithmetic (oracle_counts_to_sinogram,
pinned by closed forms in tests/test_oracle_pins.py): weights and bright field
bit-exact (the same float sums in the same order), projections within 2 ulp
(the logarithm comes from different math libraries)."""
import numpy as np
import pytest

import py_oracle
from timbir_python_amd import ingest, xt_hip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,bins,bh", [((7, 8, 24), (4, 12), 0.0), ((5, 12, 30), (3, 10), 0.35),
                                          ((3, 4, 16), (4, 16), -0.1)], ids=str)
def test_counts_to_sinogram_matches_oracle(shape, bins, bh):
    print(f"[gpu] counts to sinogram {shape} -> {bins}", flush=True)
    rng = np.random.default_rng(sum(shape))
    n_p, t_in, r_in = shape
    white = rng.integers(3000, 5000, (6, t_in, r_in)).astype(np.uint16)
    dark = rng.integers(80, 120, (5, t_in, r_in)).astype(np.uint16)
    mu = rng.uniform(0.0, 3.0, shape)
    counts = (100 + 4000 * np.exp(-mu) + rng.normal(0, 20, shape)).clip(0, 65535).astype(np.uint16)
    N_t, N_r = bins
    p, w, b = xt_hip.counts_to_sinogram(counts, white, dark, N_t, N_r, bh)
    po, wo, bo = py_oracle.Oracle(False).counts_to_sinogram(counts, white, dark, N_t, N_r, bh)
    np.testing.assert_array_equal(w, wo)
    np.testing.assert_array_equal(b, bo)
    ulp = np.abs(p.view(np.int32).astype(np.int64) - po.view(np.int32).astype(np.int64))
    print(f"  projection: max {ulp.max()} ulp, {np.count_nonzero(ulp)} of {ulp.size} differ")
    assert ulp.max() <= 2


def test_roi_and_run_folder(tmp_path):
    """roi() takes the reference's region (flat frames without the first and
    last, centred columns) and write_run_folder writes the engine's inputs."""
    rng = np.random.default_rng(3)
    data = rng.integers(500, 4000, (9, 10, 37)).astype(np.uint16)
    white = rng.integers(4000, 4200, (5, 10, 37)).astype(np.uint16)
    dark = rng.integers(90, 110, (4, 10, 37)).astype(np.uint16)
    counts, w, d, true_len = ingest.roi(data, white, dark, N_p=6, proj_start=2, slice_begin=2,
                                        slice_num=8, x_start=1, x_width=34, N_r=8)
    assert true_len == 32 and counts.shape == (6, 8, 32) and w.shape == (3, 8, 32)
    np.testing.assert_array_equal(counts, data[2:8, 2:10, 2:34])
    np.testing.assert_array_equal(d, dark[1:3, 2:10, 2:34])
    proj, wt, bright = ingest.write_run_folder(str(tmp_path), counts, w, d, N_t=4, N_r=8, bh=0.2)
    got = np.fromfile(tmp_path / "projection_n0.bin", np.float32).reshape(6, 8, 4)
    np.testing.assert_array_equal(got, proj)
    assert np.isfinite(proj).all() and (wt > 0).all()
