import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def oracle_dataset(ds, precision_model="f64"):
    """Attach oracle-built LinkMaps and oracle-model observations to a synth Dataset."""
    from oracle import oracle as O
    from updc_amd import synth
    lms = [O.build_linkmap(ds.vertices(k), ds.W, ds.H) for k in range(len(ds.geoms))]
    ex = [dict(sci=np.zeros((g.ih, g.iw), np.float32), wht=np.ones((g.ih, g.iw), np.float32), psf=ds.psfs[k],
               background=0.0, pedestal=0.0, linkmap=lms[k]) for k, g in enumerate(ds.geoms)]
    ref = O.RefSolver(ds.W, ds.H, ex, precision_model)
    blur = ref.blur(ds.truth)
    model = [ref._gather64(ref.ex[k], blur[k]).reshape(g.ih, g.iw) + ds.background[k]
             for k, g in enumerate(ds.geoms)]
    synth.observe(ds, model)
    ds.linkmaps = lms
    return ds


def oracle_exposures(ds):
    return [dict(sci=ds.sci[k], wht=ds.wht[k], psf=ds.psfs[k], background=ds.background[k],
                 pedestal=ds.pedestal[k], linkmap=ds.linkmaps[k]) for k in range(len(ds.geoms))]


@pytest.fixture(scope="session")
def small_ds():
    from updc_amd import synth
    return oracle_dataset(synth.small_config())


def as_abs32(img: bytes) -> bytes:
    """The same links and fractions in the ABS32 encoding (ULM2 encoding 1,
    updc_core.c:40-54, 463-483): header, offset[nin+1], target[L], frac[L]."""
    from oracle import oracle as O
    p = O.parse_linkmap(img)
    hdr = np.frombuffer(img[:64], np.uint8).copy()
    hdr.view(np.uint32)[2] = 1  # encoding
    parts = [hdr.tobytes(), p["offset"].astype("<u4").tobytes(), p["target"].astype("<u4").tobytes(),
             p["frac"].astype("<f4").tobytes()]
    out = b"".join(parts)
    assert len(out) == 64 + 4 * (p["nin"] + 1) + 8 * p["nlinks"]
    return out
