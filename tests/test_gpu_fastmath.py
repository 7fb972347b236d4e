"""Accuracy of the FP32 fast paths of the generated kernels (mx::sqrt_f32, mx::sincos_f32,
mx::sin_f32, mx::cos_f32, mx::atan2_f32), measured on the device against FP64 over every
FP32 bit pattern (atan2: 2^33 argument pairs, random and near the octant boundaries): every
result within 4 ulp of the FP32 value nearest the FP64 result; non-finite results exact.
Also: the library form (option "libm") and the fast form of the hillshade kernel agree within
the per-node bounds."""
import numpy as np
import pytest

from map_amd import workloads as W

pytestmark = pytest.mark.gpu

BOUND = {"sqrt": 1.0, "sin": 4.0, "cos": 4.0, "sincos": 4.0, "atan2": 4.0, "atan2_near": 4.0}


@pytest.mark.parametrize("fn", ["sqrt", "sincos", "sin", "cos", "atan2", "atan2_near"])
def test_fast_path_every_fp32_input(G, fn):
    print("case math check", fn, flush=True)
    r = G.math_check(fn)
    print("  %s: %s (worst input 0x%08x = %r)" % (fn, r, r["worst_input_bits"],
                                                float(np.uint32(r["worst_input_bits"]).view(np.float32))), flush=True)
    assert r["n"] >= 1 << 32
    assert r["n_over_4ulp"] == 0 and r["max_ulp"] <= BOUND[fn], r


def test_library_and_fast_forms_agree(G, O):
    """Hillshade of a DEM with the fast paths and with the device library: slope, aspect and
    hillshade each within the bound the per-node errors allow of each other, and both as close
    to the FP64 oracle as the FP32-semantics oracle is (within twice its error)."""
    from tests.test_gpu_parity import dem_array
    dem = dem_array(600, 420, 12)
    bs = (256, 128)
    outs = {}
    try:
        for libm in (0, 1):
            G.set_option("libm", libm)
            D = G.from_array(dem, bs)
            outs[libm] = [G.to_array(f(G, D)).astype(np.float64) for f in (W.slope, W.aspect, W.hillshade)]
    finally:
        G.set_option("libm", 0)
    O.set_mode("f64")
    od = O.from_array(dem, bs)
    ref64 = [np.asarray(f(O, od).val, np.float64) for f in (W.slope, W.aspect, W.hillshade)]
    O.set_mode("f32")
    ref32 = [f(O, O.from_array(dem, bs)).val.astype(np.float64) for f in (W.slope, W.aspect, W.hillshade)]
    for k, name in enumerate(("slope", "aspect", "hillshade")):
        e_fast = np.abs(outs[0][k] - ref64[k]).max()
        e_lib = np.abs(outs[1][k] - ref64[k]).max()
        e_o = np.abs(ref32[k] - ref64[k]).max()
        print("  %s: max |fast - fp64| %.3g, |libm - fp64| %.3g, |fp32 oracle - fp64| %.3g" % (name, e_fast, e_lib, e_o),
              flush=True)
        assert e_fast <= 2 * e_o + 1e-6 and e_lib <= 2 * e_o + 1e-6
