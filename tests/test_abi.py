"""The drop-in boundary: libnova.so loads, exports every entry point the headers declare (the
75 of nova's ctypes layer plus the additive ones), the structure layouts match nova's, and the
error contracts hold. CPU only: no call here reaches the GPU."""
import ctypes as ct
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from nova_amd import _abi


def header_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        text = open(h).read()
        names.update(re.findall(r'extern\s+"C"\s+int\s+(\w+)\s*\(', text))
    return names


def test_exports_every_declared_symbol(lib):
    declared = header_symbols()
    assert len(declared) >= 75
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_abi.ABI_SIGNATURES) <= declared
    assert len(_abi.ABI_SIGNATURES) == 75
    assert _abi.missing_symbols(lib) == []


def test_struct_layouts():
    for cls, (size, offs) in _abi.LAYOUTS.items():
        assert ct.sizeof(cls) == size, cls.__name__
        for f, o in offs.items():
            assert getattr(cls, f).offset == o, (cls.__name__, f)


def test_build_info(lib):
    buf = ct.create_string_buffer(256)
    assert lib.nova_amd_build_info(buf, 256) == 0
    assert b"gfx950" in buf.value


def _mdp3():
    from nova_amd.mdp import MDP
    S = np.array([[[0, 1], [0, -1]], [[1, -1], [2, -1]], [[2, -1], [2, -1]]], np.int32)
    T = np.array([[[0.5, 0.5], [1, 0]], [[1, 0], [1, 0]], [[1, 0], [1, 0]]], np.float32)
    R = np.array([[-1, -2], [-1, -2], [0, 0]], np.float32)
    return MDP.from_arrays(S, T, R, gamma=0.9, horizon=5)


def test_error_contracts(lib):
    # mdp_vi_cpu_tests.cpp:256-336 / mdp_vi_gpu.cu:166-178 style checks, none touch the GPU
    mdp = _mdp3()
    vi = _abi.NovaMDPVICPU()
    pol = _abi.NovaMDPValueFunction()
    assert lib.mdp_vi_execute_cpu(None, ct.byref(vi), ct.byref(pol)) == 1
    assert lib.mdp_vi_execute_cpu(ct.byref(mdp), None, ct.byref(pol)) == 1
    assert lib.mdp_vi_execute_cpu(ct.byref(mdp), ct.byref(vi), None) == 1
    old = mdp.gamma
    mdp.gamma = 1.5
    assert lib.mdp_vi_execute_cpu(ct.byref(mdp), ct.byref(vi), ct.byref(pol)) == 1
    mdp.gamma = old
    vig = _abi.NovaMDPVIGPU()
    vig.numThreads = 33    # not a multiple of 32: NOVA_ERROR_INVALID_CUDA_PARAM, no device call
    assert lib.mdp_vi_execute_gpu(ct.byref(mdp), ct.byref(vig), ct.byref(pol)) == 2
    vig.numThreads = 64    # host arrays fine but nothing resident on the device
    assert lib.mdp_vi_execute_gpu(ct.byref(mdp), ct.byref(vig), ct.byref(pol)) == 1
    # model initialization rejects n < ns, s0 >= n, ng >= n (mdp_model_cpu.cpp:38-44)
    m2 = _abi.NovaMDP()
    assert lib.mdp_initialize_cpu(ct.byref(m2), 2, 3, 1, 0.9, 1, 0.0, 0, 0) == 1
    assert lib.mdp_initialize_cpu(ct.byref(m2), 3, 2, 1, 0.9, 1, 0.0, 3, 0) == 1
    assert lib.mdp_initialize_cpu(ct.byref(m2), 3, 2, 1, 0.9, 1, 0.0, 0, 3) == 1
    # value-function initialize rejects non-null arrays (mdp_value_function.cpp:36-40)
    assert lib.mdp_value_function_initialize(ct.byref(pol), 3, 2, 0) == 0
    assert lib.mdp_value_function_initialize(ct.byref(pol), 3, 2, 0) == 1
    assert lib.mdp_value_function_uninitialize(ct.byref(pol)) == 0
    # out-of-scope entry points exist and refuse
    lao = _abi.NovaSSPLAOStarCPU()
    assert lib.ssp_lao_star_execute_cpu(ct.byref(mdp), ct.byref(lao), ct.byref(pol)) == 1


def test_initialize_twice_does_not_leak_or_crash(lib):
    """nova's Python layer initializes in the constructor and again in execute; the library
    releases the first buffers instead of leaking them, and tolerates garbage pointers."""
    mdp = _mdp3()
    vi = _abi.NovaMDPVICPU()
    assert lib.mdp_vi_initialize_cpu(ct.byref(mdp), ct.byref(vi)) == 0
    first = ct.cast(vi.V, ct.c_void_p).value
    assert lib.mdp_vi_initialize_cpu(ct.byref(mdp), ct.byref(vi)) == 0
    assert lib.mdp_vi_uninitialize_cpu(ct.byref(mdp), ct.byref(vi)) == 0
    assert first is not None
    junk = _abi.NovaMDPVICPU()
    junk.V = ct.cast(ct.c_void_p(0x1234), _abi.c_float_p)   # an uninitialised struct
    assert lib.mdp_vi_initialize_cpu(ct.byref(mdp), ct.byref(junk)) == 0
    assert lib.mdp_vi_uninitialize_cpu(ct.byref(mdp), ct.byref(junk)) == 0


def test_impossible_observation_warns():
    from nova_amd.pomdp import POMDP
    S = np.array([[[0], [0]], [[1], [1]]], np.int32)
    T = np.ones((2, 2, 1), np.float32)
    O = np.array([[[1, 0], [1, 0]], [[1, 0], [1, 0]]], np.float32)   # o = 1 never happens
    R = np.zeros((2, 2), np.float32)
    p = POMDP.from_arrays(S, T, O, R, np.zeros((1, 1), np.int32), np.ones((1, 1), np.float32),
                          0.9, 1)
    rc, _ = p.belief_update(np.array([1.0, 0.0], np.float32), 0, 1)
    assert rc == 10   # NOVA_WARNING_INVALID_BELIEF (pomdp_model_cpu.cpp:126-130)
    rc, bp = p.belief_update(np.array([1.0, 0.0], np.float32), 0, 0)
    assert rc == 0 and abs(bp.sum() - 1) < 1e-6


def test_layout_asserts_compile_against_both_header_sets():
    """nova_amd/csrc/abi_layout.cpp asserts every ABI struct's size and field offsets at
    compile time; it must compile against this tree's headers and, when the reference tree is
    present, against nova's own headers (so the asserted numbers are nova's layout)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ absent")
    src = os.path.join(ROOT, "nova_amd", "csrc", "abi_layout.cpp")
    incs = [os.path.join(ROOT, "include")]
    if os.path.isdir("/root/reference/include"):
        incs.append("/root/reference/include")
    for inc in incs:
        p = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-I" + inc, src],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert p.returncode == 0, (inc, p.stdout)
