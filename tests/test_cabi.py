"""The C-ABI library loads without a GPU and exports exactly what include/*.h declares."""

import ctypes
import glob
import os
import re
import sys

import pytest

from pymoresane_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|const char\*|void)\s+\**(pms_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for must in ("pms_init", "pms_iuwt_decompose", "pms_iuwt_recompose", "pms_fft_convolve", "pms_psf_spectrum",
                 "pms_estimate_threshold", "pms_apply_threshold", "pms_source_extraction", "pms_snr_ratio",
                 "pms_moresane", "pms_moresane_by_scale", "pms_batch_run"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert declared_functions() == set(_lib.SIGNATURES)


def test_struct_layouts_match_the_header():
    # pms_trace_event: 6 ints + 2 doubles; pms_moresane_params as declared
    assert ctypes.sizeof(_lib.TraceEvent) == 40
    assert ctypes.sizeof(_lib.MoresaneParams) == 4 * 4 + 4 * 8 + 6 * 4 + 8 + 4 * 4


def test_work_size_queries_need_no_gpu():
    lib = _lib.load()
    assert lib.pms_iuwt_work_bytes(64, 48) == 2 * 64 * 48 * 4
    assert lib.pms_fft_spectrum_bytes(512) == (512 // 2 + 1) * 512 * 8
    assert lib.pms_fft_work_bytes(256, 512) >= 2 * (257) * 256 * 8
    assert lib.pms_extract_work_bytes(3, 64, 64) > 3 * 64 * 64 * 4
    assert lib.pms_iuwt_work_bytes(-1, 5) == 0


def test_spectrum_layout_conversion_is_a_permutation():
    """The host-side rfft2 -> library layout conversion moves every bin exactly once (no GPU)."""
    import numpy as np
    from pymoresane_amd.iuwt_convolution import spectrum_positions
    for p in (4, 8, 32, 512, 8192, 6, 96, 1000, 6000):
        pos = spectrum_positions(p)
        assert sorted(pos.tolist()) == list(range(p))


def test_bad_arguments_fail_loudly_without_gpu():
    lib = _lib.load()
    out = (ctypes.c_float * 4)()
    st = lib.pms_spectrum_from_rfft2(None, 8, out)
    assert st != 0 and b"NULL" in lib.pms_last_error()
    p = 34  # 2 x 17: a prime factor above 13
    src = (ctypes.c_float * (2 * p * (p // 2 + 1)))()
    big = (ctypes.c_float * (2 * p * (p // 2 + 1)))()
    st = lib.pms_spectrum_from_rfft2(src, p, big)
    assert st == -1 and b"not supported" in lib.pms_last_error()
    st = lib.pms_spectrum_from_rfft2(src, 7, big)  # odd frame side
    assert st == -1 and b"not supported" in lib.pms_last_error()


def test_modes_are_validated_before_any_device_work():
    from pymoresane_amd import iuwt, iuwt_toolbox, iuwt_convolution
    import numpy as np
    with pytest.raises(ValueError, match="CPU path"):
        iuwt.iuwt_decomposition(np.zeros((8, 8)), 2, mode="ser")
    with pytest.raises(ValueError, match="unknown"):
        iuwt.iuwt_recomposition(np.zeros((2, 8, 8)), mode="cuda")
    with pytest.raises(ValueError):
        iuwt_convolution.fft_convolve(np.zeros((8, 8)), None, conv_device="cpu")
    with pytest.raises(ValueError):
        iuwt_toolbox.source_extraction(np.zeros((1, 8, 8)), 0.5, mode="cpu")


def test_hip_path_without_gpu_raises_instead_of_falling_back():
    if _lib.gpu_available():
        pytest.skip("a GPU is visible")
    from pymoresane_amd import iuwt
    import numpy as np
    with pytest.raises(_lib.PmsError, match="no HIP device"):
        iuwt.iuwt_decomposition(np.zeros((8, 8), np.float32), 2)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says so, before any device work."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr and res.stdout == ""
