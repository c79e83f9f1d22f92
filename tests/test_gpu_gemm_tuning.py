"""The MFMA tile-width autotuner (tsvd_gemm_tuning): while on, the first call
of each product kind runs with both tile widths and keeps the faster.  A
batch of frPCA problems (the drop-in tsvd_frpca_csr_batch, rank 128, so the
sketch, rotation and leaf products all run) gives the same factors with the
tuner off, while it tunes, and with the recorded choices (the widths only
change which workgroup computes a tile, not the order of any sum; the FP64
Grams' atomic K split is never tuned and adds in arrival order, so the
comparison is to rounding: singular values within 1e-9 relative)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def test_tile_tuning_records_choices_and_keeps_results(capi):
    rng = np.random.default_rng(11)
    mats = [sp.random(3000, 2500, density=0.02, random_state=int(rng.integers(1 << 30)), format="csr") for _ in range(3)]
    ref = capi.frpca_csr_batch(mats, 128, 6, seed=3)
    capi.gemm_tuning(True)
    try:
        tuned = capi.frpca_csr_batch(mats, 128, 6, seed=3)
    finally:
        capi.gemm_tuning(False)
    rec = capi.gemm_tuning_read()
    assert rec, "no product kind was tuned"
    for key, r in rec.items():
        assert r["tile_n"] in (80, 144), key
        assert r["ms_tile80"] > 0 and r["ms_tile144"] > 0, key
    again = capi.frpca_csr_batch(mats, 128, 6, seed=3)
    for (U0, S0, _), (U1, S1, _), (U2, S2, _) in zip(ref, tuned, again):
        for S in (S1, S2):
            assert np.max(np.abs(S - S0) / np.abs(S0).max()) < 1e-9
        for U, S in ((U1, S1), (U2, S2)):
            assert np.abs((U * S) @ (U * S).T - (U0 * S0) @ (U0 * S0).T).max() < 1e-8 * S0.max() ** 2
    print({k.split("|")[0]: (v["tile_n"], round(v["ms_tile80"], 3), round(v["ms_tile144"], 3)) for k, v in rec.items()})
