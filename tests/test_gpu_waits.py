"""Bounded waits of the stage-1 kernel: the host forces every wait to give up
(a spin limit of 0 polls) and checks that the kernel still ends, that the
failure is reported as PF_EHIP, and that the next launch with the normal
limit is exact again.  Every kernel exits normally in this test: a wave that
gives up finishes its rows without waiting."""
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


def _edges(ws, plan):
    n1, n2 = plan.grids[0]
    L = plan.L
    return (ws["edge_D"].cpu().numpy().reshape(n1, n2, 2, L),
            ws["edge_P"].cpu().numpy().reshape(n1, n2, 2, L))


@pytest.mark.parametrize("dtype,L,cfg", [("float64", 300, (2, 2, 16)),
                                         ("float32", 200, (2, 1, 8))])
def test_wait_gives_up_reports_and_recovers(torch_gpu, oracle_mod, dtype, L, cfg):
    """Waits between waves of a workgroup and between strips of a tile; the
    FP32 case keeps tiles of at most 200 rows of integer costs, where the
    FP32 DP equals the FP64 oracle exactly."""
    torch = torch_gpu
    from parflexdtw_amd import _lib, engine
    from parflexdtw_amd._plan import BatchPlan
    rng = np.random.RandomState(5)
    C = rng.randint(0, 8, size=(333, 410)).astype(np.float64)
    plan = BatchPlan([C.shape], L)
    sw = engine.StepWeights(STEPS, [2, 3, 3])
    C_dev = torch.from_numpy(C).to(getattr(torch, dtype)).cuda()
    ws = plan.allocate(C_dev.device)
    lib = _lib.lib()
    prev = ctypes.c_int64()
    assert lib.pf_set_wait_limit(0, ctypes.byref(prev)) == 0
    try:
        assert prev.value > 1000
        engine.stage1(C_dev, plan, sw, ws, config=cfg)
        with pytest.raises(_lib.ParflexError, match="gave up"):
            engine.device_status()
        engine.device_status()  # the word is cleared
    finally:
        assert lib.pf_set_wait_limit(prev.value, None) == 0
    engine.stage1(C_dev, plan, sw, ws, config=cfg)
    engine.device_status()
    eD, eP = _edges(ws, plan)
    s1 = oracle_mod.run_stage1(C, L, STEPS, [2, 3, 3])
    n1, n2 = plan.grids[0]
    for i in range(n1):
        for j in range(n2):
            t = plan.tiles[i * n2 + j]
            R, W = int(t[2]), int(t[3])
            assert np.array_equal(eD[i, j, 0, :W], s1["edge_D"][i, j, 0, :W])
            assert np.array_equal(eP[i, j, 1, :R], s1["edge_P"][i, j, 1, :R])


@pytest.mark.parametrize("cfg", [(1, 4, 16), (1, 8, 64)])
def test_sweep4_wait_gives_up_reports_and_recovers(torch_gpu, cfg):
    """The features-form sweep with 4 columns per lane (k_sweep4): its ring
    and progress waits give up at a limit of 0 polls, the launch ends and
    reports PF_EHIP; with the normal limit its edges and back-pointers
    equal the materialised BF16 form's again."""
    torch = torch_gpu
    from parflexdtw_amd import _lib, engine
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.synth import make_pair
    F1, F2, _ = make_pair(700, 760, seed=31)
    L = 1000 if cfg[1] == 4 else 2000
    fus = BatchRunner([(F1, F2)], L, STEPS, [2, 3, 3], beta=0.1, dtype=torch.bfloat16,
                      fused=True)
    fus.cost()
    lib = _lib.lib()
    prev = ctypes.c_int64()
    assert lib.pf_set_wait_limit(0, ctypes.byref(prev)) == 0
    try:
        engine.stage1(fus.C, fus.plan, fus.sw, fus.ws, config=cfg, feat=fus.feat)
        with pytest.raises(_lib.ParflexError, match="gave up"):
            engine.device_status()
        engine.device_status()
    finally:
        assert lib.pf_set_wait_limit(prev.value, None) == 0
    engine.stage1(fus.C, fus.plan, fus.sw, fus.ws, config=cfg, feat=fus.feat)
    engine.device_status()
    ref = BatchRunner([(F1, F2)], L, STEPS, [2, 3, 3], beta=0.1, dtype=torch.bfloat16)
    ref.cost()
    engine.stage1(ref.C, ref.plan, ref.sw, ref.ws)
    torch.cuda.synchronize()
    assert torch.equal(ref.ws["Bp"], fus.ws["Bp"])
    t = ref.plan.tiles[0]
    R, W = int(t[2]), int(t[3])
    for k in ("edge_D", "edge_P"):
        a = ref.ws[k].cpu().numpy().reshape(2, L)
        b = fus.ws[k].cpu().numpy().reshape(2, L)
        assert np.array_equal(a[0, :W], b[0, :W]) and np.array_equal(a[1, :R], b[1, :R]), k


def test_wait_limit_rejects_bad_values(torch_gpu):
    from parflexdtw_amd import _lib
    assert _lib.lib().pf_set_wait_limit(-1, None) == _lib.PF_EINVAL
    assert _lib.lib().pf_set_wait_limit(1 << 31, None) == _lib.PF_EINVAL


def test_stage2_dag_wait_gives_up_reports_and_recovers(torch_gpu):
    """The stage-2 tile DAG (pf_stage2_propagate_dag): at a limit of 0 polls
    a tile whose neighbour is not done gives up at once (a 12 x 14 grid of
    live tiles: tiles far from the origin take tickets long before their
    neighbours finish), the launch ends and reports PF_EHIP; with the normal
    limit its results equal the full stage 2's again."""
    torch = torch_gpu
    from parflexdtw_amd import _lib, engine
    from parflexdtw_amd.engine import BatchRunner
    from parflexdtw_amd.synth import make_pair
    F1, F2, _ = make_pair(2200, 2600, seed=33)
    L = 200
    br = BatchRunner([(F1, F2)], L, STEPS, [2, 3, 3], beta=0.1, dtype=torch.bfloat16)
    br.launch()
    torch.cuda.synchronize()
    plan, ws = br.plan, br.ws
    live = np.ones(plan.ntiles, bool)
    ref = dict(ws)
    for k in ("D_arr", "L_arr", "best_f", "best_i", "path", "path_len"):
        ref[k] = ws[k].clone()
    lib = _lib.lib()
    prev = ctypes.c_int64()
    assert lib.pf_set_wait_limit(0, ctypes.byref(prev)) == 0
    try:
        keep = engine.stage2_listed(plan, br.sw, ws, live, dag=True)
        with pytest.raises(_lib.ParflexError, match="gave up"):
            engine.device_status()
        engine.device_status()
    finally:
        assert lib.pf_set_wait_limit(prev.value, None) == 0
    keep = engine.stage2_listed(plan, br.sw, ws, live, dag=True)
    torch.cuda.synchronize()
    del keep
    engine.device_status()
    for k in ("D_arr", "best_f", "best_i", "path_len"):
        x, y = ref[k], ws[k]
        if x.dtype == torch.float64:
            x, y = x.view(torch.int64), y.view(torch.int64)
        assert torch.equal(x, y), k
    ra, rb = engine.collect(plan, ref), engine.collect(plan, ws)
    assert np.array_equal(ra.paths[0], rb.paths[0])
