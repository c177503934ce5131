"""Host-side logic of the learned selector (no GPU): the class table, the
merge of generated parts, the policy costs and the evaluation arithmetic."""
import numpy as np

from scalapack_amd import selector as S


def _part(n, split, seed):
    rng = np.random.default_rng(seed)
    return dict(features=rng.normal(size=(n, S.NIN)).astype(np.float32),
                labels=rng.integers(0, S.NOUT, n).astype(np.int32),
                times=rng.uniform(0.1, 1.0, (n, S.NOUT)), t_default=rng.uniform(0.5, 2.0, n),
                t_features=np.full(n, 0.01), meta=np.zeros((n, 3)),
                iters=np.zeros((n, S.NOUT), np.int32), sresid=np.zeros((n, S.NOUT)),
                split=np.array(split), start=np.array(0))


def test_choice_table_matches_the_network():
    assert len(S.CHOICES) == S.NOUT == 7
    assert S.CHOICES[S.FP64_CLASS][0] == "fp64"
    # every low-precision factorization appears with both refinements
    for prec in ("fp8", "bf16x3", "bf16x6"):
        assert (prec, "ir") in S.CHOICES and (prec, "gmres") in S.CHOICES
    assert S.NPARAM == 5471
    assert S.init_params(np.zeros(12), np.ones(12)).size == S.NPARAM


def test_sizes_are_log_uniform_over_config_range():
    assert S.SIZES[0] == 1024 and S.SIZES[-1] == 32768
    r = np.diff(np.log2(S.SIZES))
    assert np.allclose(r, 0.5, atol=0.01)


def test_merge_puts_training_first():
    tr = [_part(5, "train", 1), _part(3, "train", 2)]
    ho = [_part(4, "heldout", 3)]
    d = S.merge(tr, ho)
    assert int(d["n_train"]) == 8 and d["labels"].shape == (12,)
    assert np.array_equal(d["features"][8:], ho[0]["features"])


def test_evaluate_costs_follow_the_recorded_times():
    tr, ho = _part(6, "train", 4), _part(5, "heldout", 5)
    d = S.merge([tr], [ho])
    T = ho["times"]
    # choosing the labels reproduces t_best; choosing FP64 everywhere t_always_fp64
    ev = S.evaluate(d, ho["labels"])
    assert np.isclose(ev["t_best"], T[np.arange(5), ho["labels"]].sum())
    assert np.isclose(ev["t_network"], ev["t_best"] + ho["t_features"].sum())
    assert ev["accuracy"] == 1.0
    ev64 = S.evaluate(d, np.full(5, S.FP64_CLASS))
    assert np.isclose(ev64["t_always_fp64"], T[:, S.FP64_CLASS].sum())
    assert np.isclose(ev64["gain_vs_fp64"], ev64["t_always_fp64"] / ev64["t_network"])
    assert np.isclose(ev["t_default_policy"], ho["t_default"].sum())


def test_class_weights_equalise_classes_with_unit_sample_mean():
    y = np.array([0, 0, 1, 6, 6, 6])
    w = S.class_weights(y)
    assert np.isclose(w[y].mean(), 1.0) and w[2] == 0.0
    cnt = np.bincount(y, minlength=S.NOUT)
    tot = cnt * w
    assert np.allclose(tot[cnt > 0], tot[cnt > 0][0])
