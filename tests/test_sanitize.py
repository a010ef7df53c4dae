"""CPU sanitizer runs (AddressSanitizer + UndefinedBehaviorSanitizer, no GPU):
the library's host code driven through every argument-checking path of the
C ABI, and the C oracle on small transforms.  Built by tests/sanitize/Makefile
(host side only for the library: its kernel objects come from the normal build
in fft_onchip_amd/_build)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(ROOT, "fft_onchip_amd", "_build", "sanitize")
KERNEL_OBJS = [os.path.join(ROOT, "fft_onchip_amd", "_build", f"{k}.o")
               for k in ("radix_kernels", "mfma_kernels", "mfma_prec_kernels", "batch_kernels",
                             "device_tables")]

pytestmark = pytest.mark.skipif(not os.path.isdir(SAN), reason="sanitizer sources not present")


@pytest.fixture(scope="module")
def built():
    if not all(os.path.exists(p) for p in KERNEL_OBJS):
        pytest.skip("library not built (make -C fft_onchip_amd/csrc)")
    subprocess.run(["make", "-j8", "-C", SAN], check=True, stdout=subprocess.DEVNULL,
                   timeout=900)
    return OUT


def _run(cmd, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


def test_abi_argument_paths_under_asan_ubsan(built, tmp_path):
    out = _run([os.path.join(built, "abi_asan"), str(tmp_path)], tmp_path)
    assert "checks passed" in out


def test_oracle_under_asan_ubsan(built, tmp_path):
    out = _run([os.path.join(built, "oracle_asan")], tmp_path)
    assert "checks passed" in out
