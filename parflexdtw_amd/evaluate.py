"""Alignment output format and evaluation: the step after alignment in the
reference's benchmark pipeline (04_Align_Benchmarks.ipynb writes one pickled
warping path per query; eval_tools.py and 05_Evaluate_Benchmarks_L.ipynb turn
paths and ground-truth beat annotations into errors and error-rate-versus-
tolerance curves).

Wire format (04_Align_Benchmarks.ipynb#cell13,15): one file
`<outdir>/<benchmark>/<system>/<queryid>.pkl` per query holding the pickled
int64 numpy array wp of shape (2, n) (row 0: query frames, row 1: reference
frames, ascending), or a pickled None for a query without an alignment.

Host functions keep the reference's names, arguments and results, so a 05-
style script runs unchanged on top of them:
  read_start_times, getGroundTruthTimestamps, eval_alignment_single,
  eval_alignment_batch                       (eval_tools.py:5-70)
  eval_file, extractPieceName, getTimestamps, loadAlignment, eval_dir,
  calc_error_rates, calc_error_rates_batch   (05_Evaluate_Benchmarks_L.ipynb)
  get_outfile, save_alignment                (04_Align_Benchmarks.ipynb)
Errors are returned as float64 numpy arrays (the notebook's eval_file returns
a pandas Series of the same values).

The batched device form evaluates many paths in one launch (pf_eval_paths):
`evaluate_paths` takes host (2, n) paths, `evaluate_batch` reads the paths a
BatchRunner left on the device without copying them back.  Its errors equal
the host evaluator's bit for bit, and its error rates equal calc_error_rates'.
"""
from __future__ import annotations

import os
import pickle
from pathlib import Path
from typing import NamedTuple

import numpy as np

from ._lib import check, lib

__all__ = ["HOP_SEC", "get_outfile", "save_alignment", "loadAlignment", "load_alignment",
           "read_start_times", "getGroundTruthTimestamps", "eval_alignment_single",
           "eval_alignment_batch", "extractPieceName", "getTimestamps", "eval_file",
           "eval_dir", "calc_error_rates", "calc_error_rates_batch", "error_rates",
           "tolerances", "rates_from_hist", "EvalResult", "evaluate_paths", "evaluate_batch",
           "align_system_gpu", "align_batch_to_files"]

# 512-sample hop at 22050 Hz (05_Evaluate_Benchmarks_L.ipynb#cell14).
HOP_SEC = 512 / 22050


# --------------------------------------------------------------------------
# wire format
# --------------------------------------------------------------------------
def get_outfile(outdir, benchmark, system, queryid, L=None):
    """<outdir>/<benchmark>/<system>[/L_<L> for sparse_parflex]/<queryid>.pkl,
    directories created (04_Align_Benchmarks.ipynb#cell13)."""
    outpath = Path(outdir) / benchmark / system
    if system == "sparse_parflex" and L is not None:
        outpath = outpath / f"L_{L}"
    outpath.mkdir(parents=True, exist_ok=True)
    return (outpath / queryid).with_suffix(".pkl")


def save_alignment(wp, outfile):
    """Write one query's alignment as the pickled (2, n) int64 array
    (04_Align_Benchmarks.ipynb#cell15: `pickle.dump(wp, ...)` when wp is not
    None).  Returns False (and writes nothing) for wp None."""
    if wp is None:
        return False
    wp = np.asarray(wp)
    if wp.ndim != 2 or wp.shape[0] != 2:
        raise ValueError(f"a warping path has shape (2, n), got {wp.shape}")
    with open(outfile, "wb") as f:
        pickle.dump(wp, f)
    return True


class _ArrayUnpickler(pickle.Unpickler):
    """Unpickler that resolves only the names a pickled numpy array (or None)
    needs, so that a hypothesis file cannot run code when it is read."""
    _ALLOWED = {
        ("numpy", "ndarray"), ("numpy", "dtype"),
        ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
        ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
        ("numpy.core.numeric", "_frombuffer"), ("numpy._core.numeric", "_frombuffer"),
        ("_codecs", "encode"),
    }

    def find_class(self, module, name):
        if (module, name) not in self._ALLOWED:
            raise pickle.UnpicklingError(f"hypothesis files hold numpy arrays only, "
                                         f"not {module}.{name}")
        return super().find_class(module, name)


def loadAlignment(hypfile):
    """Read one alignment file (05_Evaluate_Benchmarks_L.ipynb#cell11): the
    pickled (2, n) array, or None.  Only numpy arrays are accepted."""
    with open(hypfile, "rb") as f:
        return _ArrayUnpickler(f).load()


load_alignment = loadAlignment


# --------------------------------------------------------------------------
# ground truth
# --------------------------------------------------------------------------
def read_start_times(annotfile):
    """First field of every line of a .beats file that is neither empty nor a
    '%' comment and parses as a float (eval_tools.py:5-21)."""
    out = []
    with open(annotfile, "r") as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("%"):
                continue
            tok = s.split()
            try:
                out.append(float(tok[0]))
            except ValueError:
                pass
    return np.array(out)


def getGroundTruthTimestamps(query_annot_file, ref_annot_file):
    """(n, 2) array of (query, reference) start times, line by line
    (eval_tools.py:23-31); both files must hold the same number of beats."""
    q = read_start_times(query_annot_file)
    r = read_start_times(ref_annot_file)
    return np.stack([q, r], axis=1)


def _read_beat_table(path):
    import pandas as pd
    return pd.read_csv(path, header=None, sep=r"\s+", skiprows=3)


def getTimestamps(annotfile1, annotfile2):
    """Beat times present in both .beat files (three header lines, then
    `start end label` rows), paired by label: an inner join on the label
    column in the first file's order, keeping labels whose last character is
    a decimal digit (05_Evaluate_Benchmarks_L.ipynb#cell10).  Returns
    (query times, reference times) as float64 arrays."""
    a = _read_beat_table(annotfile1)
    b = _read_beat_table(annotfile2)
    m = a.merge(b, on=2, how="inner", suffixes=("_q", "_r"))
    keep = m[2].astype(str).map(lambda s: len(s) > 0 and s[-1].isdecimal()).to_numpy(bool)
    return (m["0_q"].to_numpy()[keep].astype(np.float64),
            m["0_r"].to_numpy()[keep].astype(np.float64))


def extractPieceName(fullpath):
    """'Chopin_Op068No3_Sztompka-1959_pid9170b-21' -> 'Chopin_Op068No3'
    (05_Evaluate_Benchmarks_L.ipynb#cell9)."""
    return "_".join(os.path.basename(fullpath).split("_")[0:2])


# --------------------------------------------------------------------------
# host evaluation
# --------------------------------------------------------------------------
def _path_errors(wp, gt_x, gt_y, hop):
    """np.interp(gt_x, wp[0]*hop, wp[1]*hop) - gt_y."""
    wp = np.asarray(wp)
    pred = np.interp(gt_x, wp[0, :] * hop, wp[1, :] * hop)
    return pred - np.asarray(gt_y, dtype=np.float64)


def eval_alignment_single(hypfile, query_annot_file, ref_annot_file, outfile=None):
    """Errors of the .npy path in `hypfile` (frames) at the ground-truth
    query times (eval_tools.py:33-46): np.interp(gt_q, wp[0], wp[1]) - gt_r.
    (None, None) when there is no ground truth or no hypothesis file."""
    gt = getGroundTruthTimestamps(query_annot_file, ref_annot_file)
    if gt.shape[0] == 0:
        print(f"No measures to evaluate in {hypfile}")
        return None, None
    if not os.path.exists(hypfile):
        print(f"{hypfile} does not exist")
        return None, None
    hyp = np.load(hypfile)
    return _path_errors(hyp, gt[:, 0], gt[:, 1], 1.0)


def eval_alignment_batch(exp_dir, scenarios_dir, out_dir, tsm=False, lag=0, hypFileExt=""):
    """eval_alignment_single over every scenario directory, the dict
    {scenario: errors} pickled to <out_dir>/errs.pkl (eval_tools.py:48-70)."""
    d = {}
    for sid in os.listdir(scenarios_dir):
        if tsm:
            name = "tsm" if lag == 0 else f"tsm_lag{lag}"
        else:
            name = "hyp"
        hyp = f"{exp_dir}/{sid}/{name}{hypFileExt}.npy"
        d[sid] = eval_alignment_single(hyp, f"{scenarios_dir}/{sid}/query.beats",
                                       f"{scenarios_dir}/{sid}/ref.beats")
    os.makedirs(out_dir, exist_ok=True)
    with open(f"{out_dir}/errs.pkl", "wb") as f:
        pickle.dump(d, f)
    return d


def eval_file(hypfile, annot_root_1, annot_root_2, hop_sec):
    """Errors (seconds) of one alignment file (05_Evaluate_Benchmarks_L.ipynb
    #cell8).  The file name is `<query>__<ref>.pkl`; the annotations are
    <root>/<piece>/<query or ref>.beat with piece = extractPieceName(query).
    None when the annotations or the file cannot be read; +inf per beat for
    a file holding None."""
    parts = os.path.basename(str(hypfile)).split("__")
    assert len(parts) == 2
    piece = extractPieceName(parts[0])
    a1 = (Path(annot_root_1) / piece / parts[0]).with_suffix(".beat")
    a2 = (Path(annot_root_2) / piece / parts[1]).with_suffix(".beat")
    try:
        gt1, gt2 = getTimestamps(a1, a2)
        hyp = loadAlignment(hypfile)
    except Exception:
        print(f"Skipping hypothesis file {hypfile}")
        return None
    if hyp is None:
        return np.full(gt1.shape, np.inf)
    return _path_errors(hyp, gt1, gt2, hop_sec)


def eval_dir(hypdir, querylist, annot_root_1, annot_root_2, hop_sec, savefile=None):
    """eval_file for every `<query> <ref>` line of `querylist` whose
    <hypdir>/<query>__<ref>.pkl exists; {basename: errors}, optionally
    pickled to `savefile` (05_Evaluate_Benchmarks_L.ipynb#cell7)."""
    errs = {}
    with open(querylist, "r") as f:
        for line in f:
            parts = line.strip().split()
            assert len(parts) == 2
            base = os.path.basename(parts[0]) + "__" + os.path.basename(parts[1])
            hyp = f"{hypdir}/{base}.pkl"
            if not os.path.exists(hyp):
                continue
            e = eval_file(hyp, annot_root_1, annot_root_2, hop_sec)
            if e is not None:
                errs[base] = e
    if savefile:
        with open(savefile, "wb") as f:
            pickle.dump(errs, f)
    return errs


def tolerances(maxTol):
    """Thresholds i / 1000 s for i = 0..maxTol (one per millisecond)."""
    return np.arange(maxTol + 1, dtype=np.int64) / 1000


def error_rates(errs, maxTol):
    """Fraction of errors with |e| > i/1000 for i = 0..maxTol over all
    queries' errors (the computation of 05_Evaluate_Benchmarks_L.ipynb
    #cell23 on an in-memory dict or list).  Returns (rates, flat errors)."""
    vals = errs.values() if isinstance(errs, dict) else errs
    flat = np.concatenate([np.asarray(e, dtype=np.float64) for e in vals])
    a = np.abs(flat)
    rates = np.array([np.mean(a > t) for t in tolerances(maxTol)])
    return rates, flat


def calc_error_rates(errFile, maxTol):
    """error_rates of the {query: errors} dict pickled in `errFile`
    (05_Evaluate_Benchmarks_L.ipynb#cell23).  The file must be one this
    module (or the reference's eval_dir) wrote: it is a general pickle."""
    with open(errFile, "rb") as f:
        errs = pickle.load(f)
    return error_rates(errs, maxTol)


def calc_error_rates_batch(indir, basenames, maxTol):
    """calc_error_rates of <indir>/<basename>.pkl for every basename:
    ((len(basenames), maxTol+1) rates, [flat errors])
    (05_Evaluate_Benchmarks_L.ipynb#cell24)."""
    rates = np.zeros((len(basenames), maxTol + 1))
    vals = []
    for i, b in enumerate(basenames):
        rates[i], e = calc_error_rates(f"{indir}/{b}.pkl", maxTol)
        vals.append(e)
    return rates, vals


def rates_from_hist(hist, n):
    """Error rates from pf_eval_paths' histogram: rate[i] = #{|e| > tol[i]} / n
    with #{|e| > tol[i]} = sum_{k > i} hist[k] (exact integer counts, one
    division each, as np.mean of a boolean array)."""
    hist = np.asarray(hist, dtype=np.int64)
    above = np.cumsum(hist[::-1])[::-1]          # above[k] = sum_{k' >= k} hist[k']
    counts = above[1:]
    if n == 0:
        return np.full(len(counts), np.nan)
    return np.array([int(c) / n for c in counts])


# --------------------------------------------------------------------------
# batched device evaluation
# --------------------------------------------------------------------------
class EvalResult(NamedTuple):
    errs: list            # per query float64 errors (seconds), beats in input order
    rates: np.ndarray     # (maxTol+1,) error rate at tolerance i ms
    hist: np.ndarray      # (maxTol+2,) int64: beats with exactly k thresholds below |e|
    beats: int


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the batched evaluator needs a ROCm GPU")
    return torch


def _launch(path_t, stride, offs, lens_t, gts, hop, maxTol, stream, events=None):
    torch = _torch()
    dev = path_t.device
    nq = len(gts)
    if nq > 65535:
        raise ValueError("at most 65535 queries per call")
    gx = [np.ascontiguousarray(g[0], dtype=np.float64).reshape(-1) for g in gts]
    gy = [np.ascontiguousarray(g[1], dtype=np.float64).reshape(-1) for g in gts]
    for q, (x, y) in enumerate(zip(gx, gy)):
        if x.shape != y.shape:
            raise ValueError(f"query {q}: query and reference beat counts differ")
    cnt = np.array([len(x) for x in gx], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    total = int(off[-1])
    tol = tolerances(maxTol)
    if not 1 <= len(tol) <= 4096:
        raise ValueError("maxTol must be in [0, 4095]")
    st = stream or torch.cuda.current_stream()
    with torch.cuda.stream(st):
        offs_d = torch.from_numpy(np.ascontiguousarray(offs, dtype=np.int64)).to(dev)
        off_d = torch.from_numpy(off).to(dev)
        x_d = torch.from_numpy(np.concatenate(gx) if total else np.zeros(1)).to(dev)
        y_d = torch.from_numpy(np.concatenate(gy) if total else np.zeros(1)).to(dev)
        tol_d = torch.from_numpy(tol).to(dev)
        err_d = torch.empty(max(total, 1), dtype=torch.float64, device=dev)
        hist_d = torch.zeros(len(tol) + 1, dtype=torch.int64, device=dev)
        if events is not None:
            events[0].record(st)
        check(lib().pf_eval_paths(path_t.data_ptr(), stride, offs_d.data_ptr(), lens_t.data_ptr(),
                                  nq, off_d.data_ptr(), int(cnt.max()) if nq else 0,
                                  x_d.data_ptr(), y_d.data_ptr(), float(hop), tol_d.data_ptr(),
                                  len(tol), err_d.data_ptr(), hist_d.data_ptr(), st.cuda_stream),
              "pf_eval_paths")
        if events is not None:
            events[1].record(st)
        err = err_d.cpu().numpy()[:total]
        hist = hist_d.cpu().numpy()
    errs = [err[off[q]:off[q + 1]].copy() for q in range(nq)]
    return EvalResult(errs, rates_from_hist(hist, total), hist, total)


def evaluate_paths(paths, gts, hop_sec=HOP_SEC, maxTol=1000, stream=None, events=None):
    """Errors and error rates of many host paths in one launch.

    paths: (2, n) integer arrays (frames; rows non-decreasing; an empty or
    None path counts as no alignment: +inf errors).  gts: (query times,
    reference times) per path, seconds.  Equals eval_file's errors and
    error_rates' rates bit for bit.  events: optional pair of torch.cuda.Event
    recorded around the launch on `stream`."""
    torch = _torch()
    if len(paths) != len(gts):
        raise ValueError("one ground truth per path")
    flat, offs, lens, o = [], [], [], 0
    for q, wp in enumerate(paths):
        if wp is None:
            wp = np.zeros((2, 0), np.int64)
        wp = np.asarray(wp)
        if wp.ndim != 2 or wp.shape[0] != 2:
            raise ValueError(f"path {q}: shape {wp.shape}, expected (2, n)")
        if not np.issubdtype(wp.dtype, np.integer):
            raise ValueError(f"path {q}: frames must be integers")
        if wp.shape[1] > 1 and np.any(np.diff(wp[0]) < 0):
            raise ValueError(f"path {q}: query frames must be non-decreasing")
        n = wp.shape[1]
        flat.append(wp.astype(np.int64).reshape(-1))
        offs.append((o, o + n))
        lens.append(n)
        o += 2 * n
    host = np.concatenate(flat) if o else np.zeros(1, np.int64)
    path_t = torch.from_numpy(host).cuda()
    lens_t = torch.from_numpy(np.array(lens, dtype=np.int64)).cuda()
    return _launch(path_t, 1, np.array(offs, dtype=np.int64).reshape(-1), lens_t, gts, hop_sec,
                   maxTol, stream, events)


def evaluate_batch(runner, gts, hop_sec=HOP_SEC, maxTol=1000, stream=None, events=None):
    """evaluate_paths on the stitched paths a BatchRunner (or any plan and
    workspace run through stage 2) holds on the device: the paths are read
    in place, nothing but the errors and the histogram comes back.  A pair
    without an endpoint has an empty path (+inf errors)."""
    plan, ws = runner.plan, runner.ws
    if len(gts) != plan.npairs:
        raise ValueError("one ground truth per pair")
    po = plan.pairs[:, 11].astype(np.int64)
    offs = np.stack([2 * po, 2 * po + 1], axis=1).reshape(-1)
    return _launch(ws["path"], 2, offs, ws["path_len"], gts, hop_sec, maxTol, stream, events)


# --------------------------------------------------------------------------
# alignment to files
# --------------------------------------------------------------------------
STEPS = np.array([1, 1, 1, 2, 2, 1]).reshape((-1, 2))
WEIGHTS = np.array([1.25, 3.0, 3.0])


def align_system_gpu(F1, F2, outfile, steps=STEPS, weights=WEIGHTS, beta=0.1, L=None):
    """The 'gpu_parflex' system of 04_Align_Benchmarks.ipynb#cell10,15:
    C = 1 - L2norm(F1)^T L2norm(F2) (built on the GPU in FP64), FP64
    ParFlex, the path written with save_alignment.  Returns (best_cost, wp)."""
    from . import Parflex_gpu
    from .gpu_flexdtw import cost_matrix_from_features_gpu
    C = cost_matrix_from_features_gpu(F1, F2)
    best_cost, wp = Parflex_gpu.parflex(C, steps, weights, beta, L, profile_dir=None)
    save_alignment(wp, outfile)
    return best_cost, wp


def align_batch_to_files(pairs, outfiles, L=None, steps=STEPS, weights=WEIGHTS, beta=0.1):
    """Every (F1, F2) pair aligned in one batched FP64 launch sequence
    (BatchRunner) and written to its file; the pipeline of
    04_Align_Benchmarks.ipynb#cell17 for 'gpu_parflex', whose queries the
    reference runs one at a time under a lock.  A pair without an endpoint
    gets no file (as wp None in 04#cell15).  Returns the BatchResult."""
    from . import engine
    from .Parflex_gpu import DEFAULT_CHUNK_LENGTH
    torch = _torch()
    if len(pairs) != len(outfiles):
        raise ValueError("one output file per pair")
    L = DEFAULT_CHUNK_LENGTH if L is None else int(L)
    br = engine.BatchRunner(pairs, L, np.asarray(steps).reshape(-1, 2), np.asarray(weights),
                            beta=beta, dtype=torch.float64)
    br.launch()
    res = br.collect()
    for p, f in enumerate(outfiles):
        if res.found[p]:
            save_alignment(res.paths[p], f)
    return res
