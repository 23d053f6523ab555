"""FITS product assembly without astropy (updc_amd/fitsout.py) against the FITS
standard: card layout, block structure, big-endian data, binary tables, the
ones' complement checksum invariant, and the engine's product checks
(updc/_engine.py:824-861, 1421-1446).  CPU only.

astropy, the reference's writer, is absent here, so byte identity with its
output is unpinned; what is pinned is what the standard fixes and what the
engine's validator reads."""
import struct

import numpy as np
import pytest

from updc_amd import fitsout as F


def oc_sum_reference(data: bytes) -> int:
    """Ones' complement sum, one word at a time with end-around carry (an
    independent restatement for small inputs)."""
    s = 0
    for (w,) in struct.iter_unpack(">I", data):
        s += w
        if s > 0xFFFFFFFF:
            s = (s & 0xFFFFFFFF) + 1
    return s


# ---------------------------------------------------------------- cards
def test_fixed_format_cards():
    # FITS 4.0 section 4.2: logical/integer/real values right-justified to column 30
    assert F.format_card("SIMPLE", True, "")[0] == "SIMPLE  =                    T".ljust(80)
    assert F.format_card("NAXIS", 0, "")[0] == "NAXIS   =                    0".ljust(80)
    assert F.format_card("BITPIX", -32, "")[0] == "BITPIX  =                  -32".ljust(80)
    c = F.format_card("MUMARKER", 1.0e-12, "marker")[0]
    assert c[:30] == "MUMARKER=              1.0E-12" and c[30:] == " / marker".ljust(50)
    assert float(F.format_card("X", 0.1, "")[0][10:30]) == 0.1
    assert float(F.format_card("X", -1.2345678901234567e-150, "")[0][10:30]) == pytest.approx(-1.2345678901234567e-150,
                                                                                             rel=1e-12)
    # character strings: quote in column 11, at least 8 characters, quotes doubled
    assert F.format_card("EXTNAME", "SCI", "")[0] == "EXTNAME = 'SCI     '".ljust(80)
    assert F.format_card("X", "it's", "")[0].startswith("X       = 'it''s   '")
    with pytest.raises(ValueError):
        F.format_card("TOOLONGKEY", 1, "")
    with pytest.raises(ValueError):
        F.format_card("X", float("nan"), "")


def test_long_strings_continue():
    text = "a/very/long/path/" * 9 + "exposure_nrcb1_cal.fits"
    cards = F.format_card("INFILE01", text, "")
    assert len(cards) > 1 and all(len(c) == 80 for c in cards)
    assert all(c.startswith("CONTINUE  '") for c in cards[1:])
    pieces = [c[c.index("'") + 1:c.rindex("'")] for c in cards]
    assert all(p.endswith("&") for p in pieces[:-1])
    assert "".join(p.rstrip("&") for p in pieces) == text


def test_history_cards_wrap_at_72():
    cards = F.format_card("HISTORY", "x" * 100, "")
    assert [c[:8] for c in cards] == ["HISTORY "] * 2 and cards[0][8:80] == "x" * 72


# ---------------------------------------------------------------- checksums
def test_ones_complement_sum_matches_word_loop():
    rng = np.random.default_rng(1)
    for n in (4, 2880, 2880 * 3):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert F.ones_complement_sum(b) == oc_sum_reference(b)
    assert F.ones_complement_sum(b"\xff\xff\xff\xff" * 2) == 0xFFFFFFFF


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_checksum_makes_every_hdu_sum_to_all_ones(seed):
    """The convention's defining property: with CHECKSUM in place, the ones'
    complement sum of the whole HDU is -0 (0xFFFFFFFF); DATASUM is the
    decimal sum of the data unit."""
    rng = np.random.default_rng(seed)
    img = rng.normal(0, 10, (17, 23)).astype(np.float32)
    raw = F.image_hdu(img, [("EXTNAME", "SCI", ""), ("BUNIT", "MJy/sr", "")])
    assert len(raw) % F.BLOCK == 0
    assert oc_sum_reference(raw) == 0xFFFFFFFF
    h = F.read_hdus(raw)[0]
    assert int(h.header["DATASUM"]) == oc_sum_reference(raw[h.data_start:h.end])
    chk = h.header["CHECKSUM"]
    assert len(chk) == 16 and all(c.isalnum() for c in chk)


def test_checksum_detects_a_flipped_byte():
    raw = bytearray(F.image_hdu(np.ones((8, 8), np.float32), [("EXTNAME", "SCI", "")]))
    h = F.read_hdus(bytes(raw))[0]
    assert F.verify_checksums(bytes(raw), h) == (True, True)
    raw[h.data_start + 5] ^= 0x10
    assert F.verify_checksums(bytes(raw), h) == (False, False)


# ---------------------------------------------------------------- the product
def _inputs(n=3, H=30, W=40, seed=5):
    rng = np.random.default_rng(seed)
    data = rng.normal(1, 0.5, (H, W)).astype(np.float32)
    wht = rng.uniform(0, 2, (H, W)).astype(np.float32)
    target = [("WCSAXES", 2, ""), ("CRPIX1", 20.5, ""), ("CRPIX2", 15.5, ""), ("CTYPE1", "RA---TAN", ""),
              ("CTYPE2", "DEC--TAN", ""), ("BUNIT", "MJy/sr", "")]
    meta = [("ORIGIN", "UPDC SignedX v17-generic-final", ""), ("NCOMBINE", n, ""),
            ("MUMARKER", 1.0e-12, "Diagnostic mu marker for frozen detector pixels"), ("ITERMAX", 40, ""),
            ("CRPIX1", 99.0, "replaced in place by extend(update=True)"),
            ("HISTORY", "Signed x is updated multiplicatively without positivity clipping.", "")]
    metrics = [dict(source=f"/data/jw0{k}_cal.fits", background=0.1 * k, background_sigma=0.01,
                    poisson_pedestal=50.0 + k, auto_positivity_pedestal=0.0,
                    nonpositive_valid_before_pedestal=7 * k, sky_var_poisson=0.3) for k in range(n)]
    freeze = [dict(iteration=40, exposure=k + 1, n_new_frozen=k, n_total_frozen=10 * k) for k in range(n)]
    psf = [dict(anchor_x=1.0 * k, anchor_y=2.0, oversample=2, center_target_x=3.0, center_target_y=4.0,
                extent_x=5.0, extent_y=6.0, input_width=63, input_height=63, mapped_width=70, mapped_height=71,
                output_width=63, output_height=63, input_sum=1.0, output_sum=0.99999, crop_retained_flux=1.0,
                crop_discarded_flux=0.0, flux_rel_error=1e-6, overlaps=1234 + k, method="POLYGON",
                wcs_used="gwcs" if k else "affine") for k in range(n)]
    names = [f"psf_{k}.fits" for k in range(n)]
    return data, wht, target, meta, metrics, freeze, psf, names


def test_product_layout_and_contents(tmp_path):
    data, wht, target, meta, metrics, freeze, psf, names = _inputs()
    path = F.write_iteration_product(tmp_path / "out_iter40_i2d.fits", data, wht, target, meta, metrics, 40,
                                     "out_iter40_i2d.fits", "ivm", freeze, psf, names)
    raw = path.read_bytes()
    assert len(raw) % F.BLOCK == 0
    hdus = F.read_hdus(raw)
    assert [h.header.get("EXTNAME", "PRIMARY") for h in hdus] == ["PRIMARY", "SCI", "WHT", "UPDC_BKG",
                                                                   "UPDC_PSFGEOM"]
    prim, sci, wh, bkg, pg = hdus
    assert [k for k, _ in prim.cards[:4]] == ["SIMPLE", "BITPIX", "NAXIS", "EXTEND"]
    assert [k for k, _ in sci.cards[:7]] == ["XTENSION", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "PCOUNT", "GCOUNT"]
    assert sci.header["XTENSION"] == "IMAGE" and sci.header["BITPIX"] == -32
    assert (sci.header["NAXIS1"], sci.header["NAXIS2"]) == (40, 30)
    np.testing.assert_array_equal(F.image_data(raw, sci), data)
    np.testing.assert_array_equal(F.image_data(raw, wh), wht)
    # the engine's keywords (_engine.py:828-847)
    for h in (prim, sci):
        assert h.header["FILENAME"] == "out_iter40_i2d.fits" and h.header["UPDCITER"] == 40
        assert h.header["MUFRTOT"] == 30 and h.header["MUFRNEW"] == 3
    assert sci.header["SCIMODE"] == "SIGNED-X"
    assert wh.header["BUNIT"] == "(MJy/sr)^-2" and wh.header["WHTTYPE"] == "SUM(W*A/J)"
    # metadata replaces an existing keyword in place; HISTORY is appended
    keys = [k for k, _ in sci.cards]
    assert keys.count("CRPIX1") == 1 and sci.header["CRPIX1"] == 99.0
    assert keys.index("CRPIX1") < keys.index("EXTNAME") < keys.index("ORIGIN")
    assert "HISTORY" in keys and prim.header["ORIGIN"] == "UPDC SignedX v17-generic-final"
    # tables
    t = F.table_rows(raw, bkg)
    assert list(t.dtype.names) == ["EXPOSURE", "INPUTFILE", "BACKGROUND", "BKG_SIGMA", "POIS_PEDESTAL",
                                   "AUTO_PEDESTAL", "N_NONPOS", "SKY_VAR_POISSON"]
    assert list(t["EXPOSURE"]) == [1, 2, 3] and t["INPUTFILE"][2].decode() == "jw02_cal.fits"
    assert list(t["N_NONPOS"]) == [0, 7, 14] and t["POIS_PEDESTAL"][1] == 51.0
    assert bkg.header["TFORM2"] == "13A" and bkg.header["TFORM7"] == "K"
    g = F.table_rows(raw, pg)
    assert g.dtype.names[:3] == ("EXPOSURE", "PSFNAME", "ANCHOR_X") and len(g.dtype.names) == 23
    assert list(g["OVERLAPS"]) == [1234, 1235, 1236] and g["WCS_USED"][0].decode() == "affine"
    assert pg.header["TFORM23"] == "6A"
    # every HDU checks out
    for h in hdus:
        assert F.verify_checksums(raw, h) == (True, True)
    F.validate_product(path, 3)


def test_mask_weight_unit(tmp_path):
    data, wht, target, meta, metrics, freeze, psf, names = _inputs(n=1)
    raw = F.iteration_product(data, wht, target, meta, metrics, 1, "x.fits", "mask", freeze, psf, names)
    assert F.read_hdus(raw)[2].header["BUNIT"] == "1"


@pytest.mark.parametrize("defect", ["nan_sci", "negative_wht", "rows", "flip"])
def test_validator_rejects(tmp_path, defect):
    data, wht, target, meta, metrics, freeze, psf, names = _inputs()
    n = 3
    if defect == "nan_sci":
        data[3, 4] = np.nan
    if defect == "negative_wht":
        wht[0, 0] = -1.0
    if defect == "rows":
        n = 4
    path = F.write_iteration_product(tmp_path / "p.fits", data, wht, target, meta, metrics, 2, "p.fits", "ivm",
                                     freeze, psf, names)
    if defect == "flip":
        raw = bytearray(path.read_bytes())
        sci = F.read_hdus(bytes(raw))[1]
        raw[sci.data_start + 17] ^= 0x01
        path.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        F.validate_product(path, n)


def test_engine_writes_and_validates_every_saved_iteration(tmp_path):
    from updc_amd import engine as E
    data, wht, target, meta, metrics, freeze, psf, names = _inputs()
    its = {20: data, 40: data * 2}
    rows = [dict(r, iteration=it) for it in (20, 40) for r in freeze]
    res = E.CoreResult(its, wht, rows, {})
    files = E.write_iteration_products(res, tmp_path / "out", "run", target, meta, metrics, "ivm", psf, names)
    assert [f.name for f in files] == ["run_iter20_i2d.fits", "run_iter40_i2d.fits"]
    raw = files[1].read_bytes()
    np.testing.assert_array_equal(F.image_data(raw, F.read_hdus(raw)[1]), data * 2)
