"""pf_stage1_noalloc: the stage-1 entry point that allocates nothing (its
strip-exchange counters and rows live in the library's per-GPU scratch
array).  Called through the C ABI and checked against pf_stage1 with a
caller-allocated workspace, bit for bit, on two calls in a row (the second on
a smaller input), and a launch whose exchange exceeds the scratch is
rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = [(1, 1), (1, 2), (2, 1)]


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run(torch, entry, C_dev, plan, sw, cfg, dtype_code):
    """One stage-1 call through the C ABI; returns host copies of the outputs."""
    from parflexdtw_amd import _lib
    lib = _lib.lib()
    ws = plan.allocate(C_dev.device)
    for k in ("edge_D", "edge_C"):
        ws[k].fill_(-7.0)
    ws["edge_P"].fill_(-7)
    ws["Bp"].fill_(-1)
    S, nw, G = cfg
    st = torch.cuda.current_stream().cuda_stream
    args_head = (dtype_code, C_dev.data_ptr(), ws["tiles"].data_ptr(), plan.ntiles, plan.L,
                 sw.sp, sw.wp, sw.K, S, nw, G)
    args_tail = (ws["Bp"].data_ptr(), ws["edge_D"].data_ptr(), ws["edge_P"].data_ptr(),
                 ws["edge_C"].data_ptr(), st)
    if entry == "noalloc":
        rc = lib.pf_stage1_noalloc(*args_head, *args_tail)
    else:
        nf, nx = ctypes.c_int64(), ctypes.c_int64()
        assert lib.pf_stage1_workspace(dtype_code, plan.ntiles, plan.L, S, ctypes.byref(nf),
                                       ctypes.byref(nx)) == 0
        flags = torch.empty(max(nf.value, 1), dtype=torch.int32, device=C_dev.device)
        xch = torch.empty(max(-(-nx.value // 8), 1), dtype=torch.float64, device=C_dev.device)
        rc = lib.pf_stage1(*args_head, flags.data_ptr(), xch.data_ptr(), *args_tail)
    assert rc == 0, _lib.last_error()
    from parflexdtw_amd import engine
    engine.device_status()
    return {k: ws[k].cpu().numpy() for k in ("edge_D", "edge_P", "edge_C", "Bp")}


CASES = [  # (C shape, L, (strips, waves, G), dtype) -- second call smaller than the first
    ((700, 1500), 700, (4, 2, 16), "float64"),
    ((333, 410), 300, (2, 2, 16), "float64"),
]


def test_noalloc_equals_pf_stage1_two_calls(torch_gpu):
    torch = torch_gpu
    from parflexdtw_amd import engine
    from parflexdtw_amd._lib import PF_BF16, PF_F32, PF_F64
    from parflexdtw_amd._plan import BatchPlan
    codes = {"float64": PF_F64, "float32": PF_F32, "bfloat16": PF_BF16}
    sw = engine.StepWeights(STEPS, [2, 3, 3])
    for case, (shape, L, cfg, dt) in enumerate(CASES + [((500, 900), 450, (3, 2, 32),
                                                         "bfloat16")]):
        rng = np.random.RandomState(40 + case)
        C = rng.rand(*shape)
        C_dev = torch.from_numpy(C).to(getattr(torch, dt)).cuda()
        plan = BatchPlan([C.shape], L)
        want = _run(torch, "workspace", C_dev, plan, sw, cfg, codes[dt])
        got = _run(torch, "noalloc", C_dev, plan, sw, cfg, codes[dt])
        for k in want:
            assert np.array_equal(got[k], want[k], equal_nan=True), (case, k)


def test_noalloc_rejects_an_exchange_beyond_the_scratch(torch_gpu):
    from parflexdtw_amd import _lib
    from parflexdtw_amd._lib import PF_EINVAL, PF_F64
    lib = _lib.lib()
    sw_steps = np.array([1, 1, 1, 2, 2, 1], np.int32)
    sw_w = np.array([2, 3, 3], np.float64)
    # 1000 tiles x 8 strips x 4000 rows x 32 bytes: far above the scratch;
    # rejected before any pointer is used (all are null here)
    rc = lib.pf_stage1_noalloc(PF_F64, None, None, 1000, 4000, sw_steps.ctypes.data,
                               sw_w.ctypes.data, 3, 8, 4, 32, None, None, None, None, None)
    assert rc == PF_EINVAL
    assert "scratch" in _lib.last_error()
