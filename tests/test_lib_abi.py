"""The C-ABI library loads and exports every entry point include/updc_hip.h
declares (no compute call: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "updc_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(updc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("updc_all", "updc_last_error", "updc_ctx_create", "updc_iterate", "updc_prepare"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from updc_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_self_test_host_geometry():
    """updc_self_test only exercises host code (updc_core.c:1545-1564)."""
    from updc_amd import _lib
    assert _lib.load().updc_self_test(0) == 0


def test_dropin_executable_self_test():
    import subprocess
    from updc_amd import _lib
    r = subprocess.run([str(_lib.EXE_PATH), "--self-test"], capture_output=True, text=True)
    assert r.returncode == 0 and "UPDC C core self-test: OK" in r.stdout
    bad = subprocess.run([str(_lib.EXE_PATH), "a", "b"], capture_output=True, text=True)
    assert bad.returncode == 1 and bad.stderr.startswith("UPDC_ALL: usage")


def test_dropin_psf_projector_self_test():
    import subprocess
    from updc_amd import _lib
    r = subprocess.run([str(_lib.PROJECTOR_PATH), "--self-test"], capture_output=True, text=True)
    assert r.returncode == 0 and "UPDC PSF projector self-test: OK" in r.stdout
    bad = subprocess.run([str(_lib.PROJECTOR_PATH), "a"], capture_output=True, text=True)
    assert bad.returncode == 1 and bad.stderr.startswith("UPDC_PSF_PROJECT: usage")


_GOOD_ARGV = ["m.txt", "96", "48", "48", "48", "2", "0", "2", "it", "w.bin", "1", "1", "1", "0", "0", "0", "wis", "1"]


@pytest.mark.parametrize("pos,value", [(1, "0"), (2, "0"), (3, "0"), (4, "0"), (5, "0"), (7, "0"), (10, "0"),
                                       (11, "0"), (12, "0"), (17, "-1"), (1, "abc")])
def test_dropin_argument_ranges(pos, value):
    """The reference's range check on argv (updc_core.c:1294-1296) runs before
    any file or device is touched: each out-of-range value exits 1 with its
    message, from the executable and from updc_all alike."""
    import subprocess
    from updc_amd import _lib
    argv = list(_GOOD_ARGV)
    argv[pos] = value
    r = subprocess.run([str(_lib.EXE_PATH)] + argv, capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr.strip().endswith("UPDC_ALL: invalid dimensions/exposures/iterations/workers/save_every")
    lib = _lib.load()
    args = [a.encode() if i in (0, 8, 9, 16) else int(a) if a.lstrip("-").isdigit() else 0
            for i, a in enumerate(argv)]
    assert lib.updc_all(*args) != 0
    assert lib.updc_last_error().decode() == "invalid dimensions/exposures/iterations/workers/save_every"
