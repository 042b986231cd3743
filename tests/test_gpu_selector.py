"""The autotuner and the learned selector end to end on a small family of
matrices: every applicable candidate times finite, the trained network's
choices have a finite regret no worse than the fixed default on the matrices
it was trained on, and a saved model reloads to the same choices."""
import numpy as np
import pytest

from clspmv_amd import matrix as M
from clspmv_amd import selector as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dataset(ctx):
    return S.build_dataset(ctx, 24, seed=3, log2_min=10, log2_max=13)


def test_autotune_times_applicable_candidates(ctx, dataset):
    feats, times, kinds, nnz = dataset
    ok = S.applicable(feats)
    assert np.all(np.isfinite(times[:, S.DEFAULT]))
    assert np.all(np.isinf(times[~ok]))
    assert np.isfinite(times).sum() > 3 * len(times)


def test_selector_learns_and_round_trips(ctx, dataset, tmp_path):
    feats, times, kinds, nnz = dataset
    sel = S.Selector.fit(ctx, feats, times, steps=800)
    pick = sel.predict(feats, times)
    r = S.regret(times, pick)
    assert np.all(np.isfinite(r))
    assert r.mean() <= S.regret(times, np.full(len(pick), S.DEFAULT)).mean() + 1e-9
    sel.save(str(tmp_path / "m.npz"))
    again = S.Selector.load(ctx, str(tmp_path / "m.npz"))
    assert np.array_equal(again.predict(feats, times), pick)
    fmt, meth, name = sel.choose(M.gen_laplacian2d(64, 64))
    assert name in S.CANDIDATE_NAMES


def test_committed_model_regret_on_held_out_matrices(ctx):
    # the network trained on MI355X timings, re-evaluated on the committed
    # dataset's held-out split (same permutation as train_selector)
    import json
    import os
    z = np.load(os.path.join(S.MODEL_DIR, "dataset.npz"), allow_pickle=False)
    feats, times = z["features"], z["times"]
    rep = json.load(open(os.path.join(S.MODEL_DIR, "report.json")))
    n = len(feats)
    perm = np.random.default_rng(1).permutation(n)
    test = np.sort(perm[:rep["n_test"]])
    sel = S.Selector.load_default(ctx)
    r = S.regret(times[test], sel.predict(feats[test], times[test]))
    assert abs(r.mean() - rep["test"]["network"]["mean"]) < 1e-6
    assert r.mean() <= 0.10
    default = S.regret(times[test], np.full(len(test), S.DEFAULT)).mean()
    assert r.mean() < 0.5 * default
