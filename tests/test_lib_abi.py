"""The C-ABI library loads on the CPU and exports every declared symbol; host
side argument validation maps to Python exceptions.  No device work here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "parflex_hip.h")) as fp:
        text = fp.read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(pf_\w+)\(", text, flags=re.M)))


def test_library_exports_header_symbols():
    from parflexdtw_amd import _lib
    h = _lib.lib()
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(h, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert h.pf_version() == 101


def _check(steps, weights):
    from parflexdtw_amd import _lib
    s = np.ascontiguousarray(np.asarray(steps, dtype=np.int32).reshape(-1))
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    return _lib.lib().pf_check_steps(s.ctypes.data, w.ctypes.data, len(w))


def test_step_set_validation():
    from parflexdtw_amd import _lib
    assert _check([(1, 1), (1, 2), (2, 1)], [2, 3, 3]) == 0
    assert _check([(1, 1), (1, 2), (2, 1), (2, 2)], [2, 3, 3, 4]) == 0
    assert _check([(1, 0), (1, 2)], [1, 1]) == 0
    # no predecessor-complete step
    assert _check([(1, 2), (2, 1)], [3, 3]) == _lib.PF_EINVAL
    assert "predecessor" in _lib.last_error()
    # horizontal steps would make a row depend on itself
    assert _check([(1, 1), (0, 1)], [1, 1]) == _lib.PF_EINVAL
    assert _check([(1, 1), (3, 1)], [1, 1]) == _lib.PF_EINVAL
    assert _check([(1, 1)] * 5, [1] * 5) == _lib.PF_EINVAL
    assert _check([(1, 1)], [float("nan")]) == _lib.PF_EINVAL


def test_stage1_workspace_sizes():
    from parflexdtw_amd import _lib
    h = _lib.lib()
    nf, nx = ctypes.c_int64(), ctypes.c_int64()
    assert h.pf_stage1_workspace(_lib.PF_F64, 25, 4000, 16, ctypes.byref(nf),
                                 ctypes.byref(nx)) == 0
    assert nf.value == 25 * 16 + 1 and nx.value == 25 * 16 * 4000 * 32
    assert h.pf_stage1_workspace(_lib.PF_BF16, 3, 100, 2, ctypes.byref(nf),
                                 ctypes.byref(nx)) == 0
    assert nf.value == 7 and nx.value == 3 * 2 * 100 * 16
    assert h.pf_stage1_workspace(_lib.PF_F32, 3, 100, 1, ctypes.byref(nf),
                                 ctypes.byref(nx)) == 0
    assert nf.value == 0 and nx.value == 0


def test_errors_map_to_exceptions():
    from parflexdtw_amd import engine
    with pytest.raises(ValueError, match="step"):
        engine.StepWeights([(2, 1)], [3])
    with pytest.raises(ValueError, match="same length"):
        engine.StepWeights([(1, 1)], [1, 2])


def test_stage1_rejects_bad_arguments_without_device_work():
    from parflexdtw_amd import _lib
    h = _lib.lib()
    s = np.array([1, 1], np.int32)
    w = np.array([1.0])
    # zero tiles: nothing to do
    assert h.pf_stage1(0, None, None, 0, 10, s.ctypes.data, w.ctypes.data, 1, 1, 1, 16, None,
                       None, None, None, None, None, None) == 0
    # unknown dtype
    fake = ctypes.c_void_p(16)
    assert h.pf_stage1(7, fake, fake, 1, 10, s.ctypes.data, w.ctypes.data, 1, 1, 1, 16, None,
                       None, fake, fake, fake, fake, None) == _lib.PF_EINVAL
    # tile length above the supported maximum
    assert h.pf_stage1(0, fake, fake, 1, 9000, s.ctypes.data, w.ctypes.data, 1, 2, 16, 16, fake,
                       fake, fake, fake, fake, fake, None) == _lib.PF_EINVAL
    # strips too narrow to cover the tile, and strips without exchange buffers
    assert h.pf_stage1(0, fake, fake, 1, 4000, s.ctypes.data, w.ctypes.data, 1, 1, 1, 16, None,
                       None, fake, fake, fake, fake, None) == _lib.PF_EINVAL
    assert h.pf_stage1(0, fake, fake, 1, 4000, s.ctypes.data, w.ctypes.data, 1, 8, 4, 16, None,
                       None, fake, fake, fake, fake, None) == _lib.PF_EINVAL
    assert h.pf_stage1(0, fake, fake, 1, 400, s.ctypes.data, w.ctypes.data, 1, 1, 4, 65, None,
                       None, fake, fake, fake, fake, None) == _lib.PF_EINVAL
    assert h.pf_tile_full(fake, 10, 3, 3, 0, 5, s.ctypes.data, w.ctypes.data, 1, fake, fake,
                          fake, None) == _lib.PF_EINVAL
    assert "Empty chunk" in _lib.last_error()


def test_product_modules_import_without_gpu():
    from parflexdtw_amd import Parflex_gpu, gpu_flexdtw  # noqa: F401
    assert Parflex_gpu.DEFAULT_CHUNK_LENGTH == 4000


def test_cpu_only_request_is_refused():
    from parflexdtw_amd import Parflex_gpu
    C = np.random.RandomState(0).rand(10, 12)
    with pytest.raises(RuntimeError, match="GPU only"):
        Parflex_gpu.parflex(C, [(1, 1), (1, 2), (2, 1)], [2, 3, 3], 0.1, 5, use_gpu=False)
    os.environ["PARFLEX_DISABLE_GPU"] = "1"
    try:
        with pytest.raises(RuntimeError, match="GPU only"):
            Parflex_gpu.parflex(C, [(1, 1), (1, 2), (2, 1)], [2, 3, 3], 0.1, 5)
    finally:
        os.environ.pop("PARFLEX_DISABLE_GPU", None)
