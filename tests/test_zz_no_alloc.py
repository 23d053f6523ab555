"""updc_op_c2r_noalloc against updc_op_c2r through the C ABI.

The allocation-free variant runs the same kernels with a static device
workspace instead of a context buffer, so its images must equal updc_op_c2r's
bit for bit.  Two calls in a row, the second on a smaller input (a smaller FFT
grid in another context, and fewer rows and columns), check that the shared
workspace carries nothing from one call to the next.  FFT shapes with a
four-step axis run in the scratch updc_prepare reserved.  A spectrum beyond
the workspace is refused with UPDC_ERR_ARG before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _prepared(ds, precision, fft_shape, n_exp=1):
    from updc_amd.solver import Context, Exposure
    c = Context(ds.W, ds.H, precision, fft_shape=fft_shape)
    for k in range(n_exp):
        c.add_exposure(Exposure(ds.sci[k], ds.wht[k], ds.psfs[k], ds.background[k], ds.pedestal[k],
                                vertices=ds.vertices(k)))
    c.prepare()
    return c


def _both(c, spec, rows, cols, scale):
    """(updc_op_c2r image, updc_op_c2r_noalloc image) of one spectrum."""
    import torch
    from updc_amd._lib import check
    lib = c.lib
    dt = torch.float64 if spec.dtype == torch.complex128 else torch.float32
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ref = torch.full((rows, cols), float("nan"), dtype=dt, device=spec.device)
    new = torch.full((rows, cols), float("nan"), dtype=dt, device=spec.device)
    check(lib.updc_op_c2r(c.h, C.c_void_p(spec.data_ptr()), C.c_void_p(ref.data_ptr()), rows, cols, cols,
                          float(scale), st))
    check(lib.updc_op_c2r_noalloc(c.h, C.c_void_p(spec.data_ptr()), C.c_void_p(new.data_ptr()), rows, cols, cols,
                                  float(scale), st))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), new.cpu().numpy()


def _spectrum(c, seed):
    import torch
    fw, fh, nc, pitch, eb = (C.c_int() for _ in range(5))
    c.lib.updc_op_layout(c.h, C.byref(fw), C.byref(fh), C.byref(nc), C.byref(pitch), C.byref(eb))
    g = torch.Generator(device="cpu").manual_seed(seed)
    cdt = torch.complex128 if eb.value == 8 else torch.complex64
    spec = torch.randn((nc.value, pitch.value), dtype=cdt, generator=g).to("cuda")
    return spec, fh.value, fw.value


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_c2r_equals_op_c2r_twice(small_ds, prec):
    from conftest import oracle_dataset
    from updc_amd import synth
    print(f"case: op_c2r_noalloc vs op_c2r, {prec}: large then small", flush=True)
    big = _prepared(small_ds, prec, (160, 144))
    tiny = oracle_dataset(synth.small_config(n_exp=1, det=16, target=32, psf=5, seed=8))
    small = _prepared(tiny, prec, (48, 40))
    spec_b, fh_b, fw_b = _spectrum(big, 1)
    spec_s, fh_s, fw_s = _spectrum(small, 2)
    assert fh_s * fw_s < fh_b * fw_b
    ref, new = _both(big, spec_b, 96, 96, 1.0 / (fh_b * fw_b))
    assert np.isfinite(ref).all() and np.array_equal(ref, new)
    # second call: smaller grid, fewer rows and columns
    ref, new = _both(small, spec_s, 20, 27, 0.5)
    assert np.isfinite(ref).all() and np.array_equal(ref, new)
    # and back on the first context with a partial crop
    ref, new = _both(big, spec_b, 7, 150, 2.0)
    assert np.isfinite(ref).all() and np.array_equal(ref, new)
    big.close()
    small.close()


@pytest.mark.parametrize("shape", [(8400, 112), (112, 8400)])
def test_noalloc_c2r_on_four_step_shapes(small_ds, shape):
    """An FFT axis too long for one workgroup (four-step) works in the scratch
    updc_prepare reserved: the allocation-free call runs and equals
    updc_op_c2r bit for bit."""
    print(f"case: op_c2r_noalloc vs op_c2r at FFT {shape}", flush=True)
    c = _prepared(small_ds, "f32", shape)
    spec, fh, fw = _spectrum(c, 3)
    ref, new = _both(c, spec, 90, 96, 1.0 / (fh * fw))
    assert np.isfinite(ref).all() and np.array_equal(ref, new)
    c.close()


def test_noalloc_c2r_refuses_a_spectrum_beyond_the_workspace(small_ds):
    """FP64 at 5600 x 5600: 2801 columns x 5600 rows x 16 B = 251 MB > 128 MiB."""
    import torch
    from updc_amd._lib import UPDC_OK
    print("case: op_c2r_noalloc beyond the workspace", flush=True)
    c = _prepared(small_ds, "f64", (5600, 5600))
    spec = torch.zeros((2801, 5600), dtype=torch.complex128, device="cuda")
    out = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    rc = c.lib.updc_op_c2r_noalloc(c.h, C.c_void_p(spec.data_ptr()), C.c_void_p(out.data_ptr()), 8, 8, 8, 1.0,
                                   C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 1 and rc != UPDC_OK
    msg = c.lib.updc_last_error().decode()
    assert "exceeds the allocation-free workspace of 134217728 bytes" in msg, msg
    torch.cuda.synchronize()
    assert not out.any()
    c.close()
