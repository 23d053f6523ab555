"""The reference's own data files for this path, committed as fixtures in
tests/golden/ (tests/golden/make_fixtures.sh copies them):

- updc/cache_abi.json: the LinkMap ABI (magic, version 2, ABS32 = 1,
  DELTA16 = 2), read by tests/test_core_freeze.py:11-14 of the reference;
- updc/core_freeze.json: SHA-256 of the frozen v17 C core
  (tests/test_core_freeze.py:5-8).

The C ABI header, the kernels, the oracle's C restatement and the LinkMap
images the oracle builds must carry the fixture's ABI. Where the reference tree
is present, its C sources must be the frozen bytes, so the file:line citations
of the oracle and the library refer to the v17 core.
"""
import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
REFERENCE = Path("/root/reference")


def _abi():
    return json.loads((GOLDEN / "cache_abi.json").read_text())["linkmap"]


def _defines(path: Path) -> dict:
    out = {}
    for m in re.finditer(r"#define\s+(\w+)\s+(0x[0-9A-Fa-f]+|\d+)u?\b", path.read_text()):
        out[m.group(1)] = int(m.group(2), 0)
    return out


def test_linkmap_abi_constants_match_the_reference_fixture():
    abi = _abi()
    magic = int(abi["magic_hex"], 16)
    hdr = _defines(ROOT / "include" / "updc_hip.h")
    assert hdr["UPDC_LINKMAP_MAGIC"] == magic and hdr["UPDC_LINKMAP_VERSION"] == abi["abi"]
    k = _defines(ROOT / "updc_amd" / "csrc" / "kernels.hpp")
    assert k["UPDC_LM_MAGIC"] == magic and k["UPDC_LM_VERSION"] == abi["abi"]
    assert k["UPDC_LM_ABS32"] == abi["encodings"]["ABS32"] and k["UPDC_LM_DELTA16"] == abi["encodings"]["DELTA16"]
    o = _defines(ROOT / "oracle" / "updc_oracle.c")
    assert o["ORC_LM_MAGIC"] == magic and o["ORC_LM_VERSION"] == abi["abi"]
    assert o["ORC_LM_ABS32"] == abi["encodings"]["ABS32"] and o["ORC_LM_DELTA16"] == abi["encodings"]["DELTA16"]


def test_oracle_linkmaps_carry_the_fixture_abi():
    """A narrow target gives DELTA16, a target wider than 65535 ABS32
    (docs/V17_VALIDATION.md:9-18); both headers carry the fixture's values."""
    from oracle import oracle as O
    from updc_amd import synth
    abi = _abi()
    g = synth.Geometry(8, 3, 2.0, 10.0, (0.3, 0.2), None, (40.0, 3.5))
    for width, enc in ((96, "DELTA16"), (70000, "ABS32")):
        geo = g if width == 96 else synth.Geometry(8, 3, 2.0, 10.0, (0.3, 0.2), None, (100.0, 3.5))
        lm = O.parse_linkmap(O.build_linkmap(geo.vertices(), width, 8))
        assert lm["magic"] == int(abi["magic_hex"], 16)
        assert lm["version"] == abi["abi"]
        assert lm["encoding"] == abi["encodings"][enc]
        assert lm["nlinks"] > 0 and np.isfinite(lm["frac"]).all()


@pytest.mark.skipif(not (REFERENCE / "updc" / "c_core").is_dir(), reason="reference tree not present")
def test_reference_c_sources_are_the_frozen_core():
    frozen = json.loads((GOLDEN / "core_freeze.json").read_text())
    assert frozen["core_freeze"] == "v17"
    for name, digest in frozen["sha256"].items():
        data = (REFERENCE / "updc" / "c_core" / name).read_bytes()
        assert hashlib.sha256(data).hexdigest() == digest, name
