"""The host half of the library and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer: both are rebuilt with the sanitizers (nova_amd/build.py
--sanitize; oracle/Makefile target san), and the host-path tests (tests/test_cpu_path.py:
every *_cpu entry point against nova's own build; tests/test_oracle.py: the oracle against
nova's build, known answers and closed forms) run in a child process with the sanitizer
runtimes preloaded. Any report fails the run (UBSan halts on its first error). The sites this
guards are nova's own undefined behaviour that the library deviates from: the missing return
of pomdp_add_new_raw_beliefs_cpu (src/pomdp/utilities/pomdp_model_cpu.cpp:157-227) and the
uninitialised argmax of the Perseus backup (src/pomdp/algorithms/pomdp_perseus_cpu.cpp:75-117),
plus every buffer the host entry points index. CPU only."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT


def _runtime(name):
    p = subprocess.run(["g++", "-print-file-name=" + name], stdout=subprocess.PIPE, text=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_host_paths_under_asan_and_ubsan():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++/make absent")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes absent")
    sys.path.insert(0, ROOT)
    from nova_amd import build as nb
    lib = nb.build_sanitized()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD="%s %s" % (asan, ubsan),
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               NOVA_AMD_LIB=lib, NOVA_AMD_ORACLE_LIB=os.path.join(ROOT, "oracle",
                                                                  "liboracle_san.so"),
               NOVA_AMD_NO_TORCH="1")
    p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "-m", "not gpu", os.path.join(ROOT, "tests", "test_cpu_path.py"),
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    out = p.stdout
    print(out[-2000:])
    assert p.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert " passed" in out
