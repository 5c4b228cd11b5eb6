"""Alignment output format and evaluation on the host (no GPU).

Pinned by the reference's own known answers: the evaluator unit tests of
unit_test_eval.ipynb, whose annotation files and paths (cells 2-56) are
rebuilt here and whose assertions (cells 62-63) are checked as written there.
The rest is pinned by closed forms (identity paths, exact integer errors)
and by counting thresholds directly."""
import os
import pickle

import numpy as np
import pytest

HOP = 512 / 22050


def _measures():
    return [f"{i}-{j}" for i in range(100 // 3 + 1) for j in (1, 2, 3)][:100]


def _rows(starts, labels):
    return [(int(s), int(s) + 1, lab) for s, lab in zip(starts, labels)]


def _write_beat(path, rows):
    """unit_test_eval.ipynb#cell2: three '%' header lines, then tab-separated
    `start end label` rows."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("%\n%\n% start_time[sec]\tend_time[sec]\tlabel\n")
        for r in rows:
            f.write("\t".join(str(v) for v in r) + "\n")


def _annotations(benchmark, case):
    """(annot1 rows, annot2 rows) of unit_test_eval.ipynb#cell7-56."""
    m = _measures()
    t = np.arange(1, 101)
    a1 = _rows(t, m)
    shift = 2 if case == "shifted_2" else 0
    if benchmark in ("matching", "post5"):
        a2 = _rows(t + shift, m)
    elif benchmark == "subseq10":
        a2 = _rows(t + shift, m)[10:20]
    elif benchmark == "partialOverlap":
        a1 = a1[:60]
        a2 = _rows(np.arange(0, 60) + shift, m[-60:])
    elif benchmark in ("pre5", "prepost5"):
        a2 = _rows(t + 5 + shift, m)
    else:
        raise ValueError(benchmark)
    return a1, a2


BENCHMARKS = ["matching", "subseq10", "partialOverlap", "pre5", "post5", "prepost5"]


def _unit_wp(end_second=100, hop=HOP):
    """unit_test_eval.ipynb#cell3: the identity path over 0..end_second/hop
    frames (a float array, as the notebook builds it)."""
    path = np.arange(0, (end_second / hop) + 1)
    return np.array([path] * 2)


@pytest.fixture(scope="module")
def unit_tree(tmp_path_factory):
    from parflexdtw_amd.evaluate import save_alignment
    root = tmp_path_factory.mktemp("unit_tests")
    for b in BENCHMARKS:
        for case in ("exact_match", "shifted_2"):
            a1, a2 = _annotations(b, case)
            _write_beat(str(root / b / case / f"{case}_annot1.beat"), a1)
            _write_beat(str(root / b / case / f"{case}_annot2.beat"), a2)
            save_alignment(_unit_wp(), root / b / case / f"{case}_annot1__{case}_annot2.pkl")
    return root


@pytest.mark.parametrize("benchmark", BENCHMARKS)
def test_unit_eval_exact_match(unit_tree, benchmark):
    """unit_test_eval.ipynb#cell62."""
    from parflexdtw_amd.evaluate import eval_file
    hyp = unit_tree / benchmark / "exact_match" / "exact_match_annot1__exact_match_annot2.pkl"
    errs = eval_file(hyp, unit_tree / benchmark, unit_tree / benchmark, HOP)
    assert errs is not None and len(errs) > 0
    if benchmark in ("matching", "subseq10", "post5"):
        assert np.all(errs == 0)
    elif benchmark == "partialOverlap":
        assert np.all(np.abs(errs) == 41)
    else:
        assert np.all(np.abs(errs) == 5)
    assert len(errs) == {"subseq10": 10, "partialOverlap": 20}.get(benchmark, 100)


@pytest.mark.parametrize("benchmark", BENCHMARKS)
def test_unit_eval_shifted(unit_tree, benchmark):
    """unit_test_eval.ipynb#cell63."""
    from parflexdtw_amd.evaluate import eval_file
    hyp = unit_tree / benchmark / "shifted_2" / "shifted_2_annot1__shifted_2_annot2.pkl"
    errs = eval_file(hyp, unit_tree / benchmark, unit_tree / benchmark, HOP)
    if benchmark in ("matching", "subseq10", "post5"):
        assert np.all(np.abs(errs) == 2)
    elif benchmark == "partialOverlap":
        assert np.all(np.abs(errs) == 39) or np.all(np.abs(errs) == 43)
    else:
        assert np.all(np.abs(errs) == 3) or np.all(np.abs(errs) == 7)


def test_eval_dir_and_error_rates(unit_tree, tmp_path):
    """eval_dir over a query list, the pickled {query: errors} file, and the
    error-rate curve of calc_error_rates / calc_error_rates_batch: a shift
    of exactly 2 s is an error at every tolerance below 2000 ms and none
    from 2000 ms on; zero errors are errors at no tolerance."""
    from parflexdtw_amd.evaluate import calc_error_rates_batch, eval_dir
    ql = tmp_path / "queries"
    ql.write_text("shifted_2_annot1 shifted_2_annot2\nexact_match_annot1 exact_match_annot2\n"
                  "missing_annot1 missing_annot2\n")
    for case in ("shifted_2", "exact_match"):
        out = tmp_path / "eval" / case
        out.mkdir(parents=True)
        hypdir = unit_tree / "matching" / case
        errs = eval_dir(str(hypdir), ql, unit_tree / "matching", unit_tree / "matching", HOP,
                        savefile=str(out / "matching.pkl"))
        assert list(errs) == [f"{case}_annot1__{case}_annot2"]
    rates, vals = calc_error_rates_batch(str(tmp_path / "eval"),
                                         ["shifted_2/matching", "exact_match/matching"], 2500)
    assert rates.shape == (2, 2501)
    assert np.all(rates[0, :2000] == 1.0) and np.all(rates[0, 2000:] == 0.0)
    assert np.all(rates[1] == 0.0)
    assert [len(v) for v in vals] == [100, 100]


def test_wire_format_round_trip_and_none(tmp_path):
    """04_Align_Benchmarks.ipynb#cell13,15 layout and pickle; a file holding
    None evaluates to +inf per beat (05_Evaluate_Benchmarks_L.ipynb#cell8)."""
    from parflexdtw_amd.evaluate import eval_file, get_outfile, loadAlignment, save_alignment
    f = get_outfile(tmp_path, "matching", "gpu_parflex", "Chopin_Op1_a__Chopin_Op1_b")
    assert f == tmp_path / "matching" / "gpu_parflex" / "Chopin_Op1_a__Chopin_Op1_b.pkl"
    g = get_outfile(tmp_path, "matching", "sparse_parflex", "q", L=2000)
    assert g.parent.name == "L_2000" and g.parent.is_dir()
    wp = np.array([[0, 1, 3, 4], [0, 2, 3, 5]], dtype=np.int64)
    assert save_alignment(wp, f)
    back = loadAlignment(f)
    assert back.dtype == np.int64 and np.array_equal(back, wp)
    with open(f, "rb") as fp:            # the plain pickle reader sees the same array
        assert np.array_equal(pickle.load(fp), wp)
    assert not save_alignment(None, g)
    # None hypothesis
    rows = _rows(np.arange(1, 11), [str(i) for i in range(10)])
    _write_beat(str(tmp_path / "ann" / "Chopin_Op1" / "Chopin_Op1_a.beat"), rows)
    _write_beat(str(tmp_path / "ann" / "Chopin_Op1" / "Chopin_Op1_b.beat"), rows)
    with open(f, "wb") as fp:
        pickle.dump(None, fp)
    e = eval_file(f, tmp_path / "ann", tmp_path / "ann", HOP)
    assert e.shape == (10,) and np.all(np.isinf(e))
    # unreadable annotations: skipped
    assert eval_file(f, tmp_path / "nowhere", tmp_path / "ann", HOP) is None


def test_loader_refuses_non_array_pickles(tmp_path):
    import collections
    from parflexdtw_amd.evaluate import loadAlignment
    f = tmp_path / "x.pkl"
    with open(f, "wb") as fp:
        pickle.dump(collections.OrderedDict(a=1), fp)
    with pytest.raises(pickle.UnpicklingError):
        loadAlignment(f)
    for proto in range(0, pickle.HIGHEST_PROTOCOL + 1):
        wp = np.arange(10, dtype=np.int64).reshape(2, 5)
        with open(f, "wb") as fp:
            pickle.dump(wp, fp, protocol=proto)
        assert np.array_equal(loadAlignment(f), wp), proto


def test_get_timestamps_pairs_by_label_and_filters(tmp_path):
    """Inner join on the label in the first file's order; labels not ending
    in a digit are dropped; labels missing from either side are dropped."""
    from parflexdtw_amd.evaluate import getTimestamps
    a = [(1.5, 2, "1-1"), (2.5, 3, "1-2"), (3.5, 4, "1-3x"), (4.5, 5, "2-1"), (5.5, 6, "9-9")]
    b = [(10.25, 11, "2-1"), (11.25, 12, "1-1"), (12.25, 13, "1-3x"), (13.25, 14, "1-2")]
    _write_beat(str(tmp_path / "a.beat"), a)
    _write_beat(str(tmp_path / "b.beat"), b)
    g1, g2 = getTimestamps(tmp_path / "a.beat", tmp_path / "b.beat")
    assert np.array_equal(g1, [1.5, 2.5, 4.5])
    assert np.array_equal(g2, [11.25, 13.25, 10.25])


def test_eval_tools_single_and_batch(tmp_path):
    """eval_tools.py:5-70: .beats start times (comments and bad lines
    skipped), .npy paths in frames, errors np.interp(gt_q) - gt_r, the
    pickled {scenario: errors} dict."""
    from parflexdtw_amd.evaluate import eval_alignment_batch, read_start_times
    sc = tmp_path / "scen" / "s1"
    sc.mkdir(parents=True)
    (sc / "query.beats").write_text("% header\n\n0.0 x\n10.0\nnot-a-number 3\n20.0 y\n")
    (sc / "ref.beats").write_text("% header\n1.0\n21.0\n40.0\n")
    assert np.array_equal(read_start_times(sc / "query.beats"), [0.0, 10.0, 20.0])
    ex = tmp_path / "exp" / "s1"
    ex.mkdir(parents=True)
    np.save(ex / "hyp.npy", np.array([[0, 10, 20], [0, 20, 40]], dtype=np.int64))
    d = eval_alignment_batch(str(tmp_path / "exp"), str(tmp_path / "scen"), str(tmp_path / "out"))
    assert np.array_equal(d["s1"], [-1.0, -1.0, 0.0])
    with open(tmp_path / "out" / "errs.pkl", "rb") as fp:
        assert np.array_equal(pickle.load(fp)["s1"], d["s1"])


def _count_above(err, tol):
    a = np.abs(np.asarray(err, dtype=np.float64))
    return (a[:, None] > tol[None, :]).sum(axis=1)


def test_rates_from_histogram_equal_error_rates():
    """The device evaluator's histogram (k = number of thresholds below
    |e|) gives exactly the rates of the direct per-threshold mean, including
    ties with a threshold, NaN (below every threshold) and +inf."""
    from parflexdtw_amd.evaluate import error_rates, rates_from_hist, tolerances
    rng = np.random.default_rng(7)
    errs = [rng.normal(0, 0.3, 500), np.array([0.0, 0.001, -0.001, 0.0015, np.nan, np.inf,
                                               -np.inf, 0.999, 1.0, 1.0000001])]
    tol = tolerances(1000)
    assert tol[1] == 0.001 and tol[1000] == 1.0 and len(tol) == 1001
    k = np.concatenate([_count_above(e, tol) for e in errs])
    hist = np.bincount(k, minlength=len(tol) + 1)
    want, flat = error_rates(errs, 1000)
    got = rates_from_hist(hist, len(flat))
    assert np.array_equal(got, want)
    assert want[0] == np.mean(np.abs(flat) > 0)
