"""The C ABI's host code under AddressSanitizer and UBSan (no GPU).

Builds libgpk with the host passes instrumented (device code as usual) and
runs the ABI tests against it in a child process with the ASan runtime
preloaded: argument validation, error strings, workspace layout arithmetic,
the batched planner's grouping and offset checks, the no-allocation entry
point's limits.  Any heap, stack or global overflow, use after free, signed
overflow, misaligned access or invalid shift in those paths ends the child
with a sanitizer report.
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "gpjax_amd" / "csrc"))

import build as gpk_build  # noqa: E402

HOST_TESTS = ["tests/test_abi.py", "tests/test_host_api.py"]

CHILD = r"""
import sys
from pathlib import Path
sys.path.insert(0, {root!r})
import gpjax_amd._lib as L
L.LIB_PATH = Path({lib!r})
assert L.lib()._name == {lib!r}
import pytest
sys.exit(pytest.main(["-q", "-p", "no:cacheprovider", "-m", "not gpu", *{tests!r}]))
"""


@pytest.fixture(scope="module")
def sanitized_lib(tmp_path_factory):
    try:
        gpk_build.hipcc()
        runtime = gpk_build.asan_runtime()
    except (RuntimeError, OSError, subprocess.CalledProcessError) as exc:
        pytest.skip(f"no ROCm toolchain here: {exc}")
    if not runtime.exists():
        pytest.skip(f"ASan runtime {runtime} not found")
    d = tmp_path_factory.mktemp("gpk_san")
    lib = gpk_build.build(jobs=min(9, os.cpu_count() or 1), out=d / "libgpk_san.so",
                          sanitize=True, build_dir=d / "obj")
    yield lib, runtime
    shutil.rmtree(d / "obj", ignore_errors=True)


def test_library_is_instrumented(sanitized_lib):
    lib, runtime = sanitized_lib
    syms = subprocess.run(["nm", "-D", str(lib)], capture_output=True, text=True,
                          check=True).stdout
    assert "__asan_report_load" in syms
    assert "__ubsan_handle_" in syms
    needed = subprocess.run(["readelf", "-d", str(lib)], capture_output=True, text=True,
                            check=True).stdout
    assert runtime.name in needed


def test_host_abi_tests_pass_under_asan_ubsan(sanitized_lib):
    lib, runtime = sanitized_lib
    env = dict(os.environ)
    # The runtime must come first among the preloaded libraries; anything
    # already preloaded stays.
    env["LD_PRELOAD"] = ":".join(p for p in (str(runtime), env.get("LD_PRELOAD", "")) if p)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    code = CHILD.format(root=str(ROOT), lib=str(lib), tests=HOST_TESTS)
    proc = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=600)
    out = proc.stdout + proc.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert "runtime error:" not in out, out[-4000:]
    assert proc.returncode == 0, out[-4000:]
    assert " passed" in proc.stdout
