"""The drivers' file loop (treesvd_amd.drivers.run_config) on the device:
config.txt -> subset file -> one edge file per snapshot -> U_all_shots CSV
(Tree_u_NC.cpp:338-379, 507-866, 898-909)."""
import numpy as np
import pytest

from treesvd_amd import drivers, gen

pytestmark = pytest.mark.gpu


def write_dataset(tmp_path, snaps, sub):
    files = []
    (tmp_path / "subset.txt").write_text("\n".join(str(int(v)) for v in sub) + "\n")
    files.append(str(tmp_path / "subset.txt"))
    for i, (s, d) in enumerate(snaps):
        f = tmp_path / f"snap{i}.txt"
        f.write_text("".join(f"{a} {b}\n" for a, b in zip(s, d)))    # trailing newline, as the tools write
        files.append(str(f))
    (tmp_path / "config.txt").write_text("\n".join(files) + "\n")
    return str(tmp_path / "config.txt")


@pytest.mark.parametrize("lazy_start", [2, 3])
def test_run_config_matches_the_pipeline_and_writes_the_csv(capi, tmp_path, lazy_start):
    s, d, n = gen.rmat_edges(11, 9000, 5)
    snaps = gen.shuffle_split(s, d, 5, 0.7, [900, 900, 900])
    sub = gen.sample_subset(snaps[0][0], 40, 5)
    cfg = write_dataset(tmp_path, snaps, sub)
    k = 8
    out = tmp_path / "U.csv"
    # a snapshot skipped by the lazy start is only ingested, so the static
    # push must come after it (dynamic_ppr_start_iter = lazy_update_start_iter)
    U_all = drivers.run_config(cfg, str(out), n, alpha=0.2, pass_=6, residuemax=1e-6, nParts=4, hierarchy_n=2,
                               dynamic_ppr_start_iter=lazy_start, lazy_update_start_iter=lazy_start, d=k,
                               precision=capi.TSVD_FP64, seed=2)
    assert U_all.shape == (len(snaps) * sub.size, k)
    # the same snapshots straight through the pipeline (the reader keeps the
    # duplicated last edge of each file, as the reference's loop does)
    pipe = capi.Pipeline(n, False, sub, 0.2, 1e-6, 4, 2, k, 6, precision=capi.TSVD_FP64, dyn_start=lazy_start,
                         lazy_start=lazy_start, seed=2)
    for it, (a, b) in enumerate(snaps, start=1):
        a2, b2 = np.r_[a, a[-1]], np.r_[b, b[-1]]
        pipe.snapshot(a2, b2)
        rows = U_all[(it - 1) * sub.size:it * sub.size]
        if lazy_start != 100 and it < lazy_start - 1:
            assert not rows.any()
            continue
        E = pipe.embedding()
        assert E.any()
        G, H = E @ E.T, rows @ rows.T
        # two runs of the same snapshots differ only by the order of atomic
        # additions, which can tip a residue that sits at rmax either way
        assert np.abs(G - H).max() <= 1e-4 * np.abs(G).max(), f"snapshot {it}"
    text = out.read_text()
    assert not text.endswith("\n") and text.count("\n") == U_all.shape[0] - 1
    back = np.array([[float(x) for x in line.split(", ")] for line in text.split("\n")])
    assert np.allclose(back, U_all, rtol=1e-5, atol=1e-300)
