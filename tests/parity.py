"""Comparison helpers for parity tests (bit-exact integers, ULP/abs bounds for FP32)."""
import numpy as np

EPS32 = np.float32(np.finfo(np.float32).eps)


def ulp32(ref):
    """Size of one FP32 ulp at each reference value."""
    r = np.abs(np.asarray(ref, np.float64)).astype(np.float32)
    return np.spacing(r).astype(np.float64)


def fp32_report(got, ref64, scale=None, max_ulp=4.0, abs_k=4.0):
    """Per-output FP32 check against an FP64 oracle.

    A cell passes when its error is within `max_ulp` FP32 ulps of the FP64
    value, or within abs_k * eps32 * S where S is the sum of the magnitudes of
    the terms of the last combination step (`scale`), so that cancellation
    near zero does not fail correct code. Returns the report dict.
    """
    got = np.asarray(got, np.float64)
    ref64 = np.asarray(ref64, np.float64)
    err = np.abs(got - ref64)
    u = err / ulp32(ref64)
    ok = u <= max_ulp
    if scale is not None:
        ok |= err <= abs_k * float(EPS32) * np.asarray(scale, np.float64)
    return {
        "max_ulp": float(np.nanmax(u)) if u.size else 0.0,
        "max_abs": float(np.nanmax(err)) if err.size else 0.0,
        "n_over_4ulp": int(np.count_nonzero(u > max_ulp)),
        "n_fail": int(np.count_nonzero(~ok)),
        "n": int(err.size),
    }


def assert_fp32(got, ref64, scale=None, what=""):
    rep = fp32_report(got, ref64, scale)
    print("%s: %s" % (what, rep), flush=True)
    assert rep["n_fail"] == 0, "%s: %d of %d cells outside the FP32 bound: %s" % (what, rep["n_fail"], rep["n"], rep)
    return rep


def assert_exact(got, ref, what=""):
    """Equal cell for cell. A NaN matches any NaN: IEEE 754 leaves the sign and payload of the
    NaN an invalid operation returns to the platform (x86 sets the sign bit, the GPU does not)."""
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    diff = got != ref
    if np.issubdtype(got.dtype, np.floating) and np.issubdtype(ref.dtype, np.floating):
        diff &= ~(np.isnan(got) & np.isnan(ref))
    bad = np.argwhere(diff)
    assert bad.size == 0, "%s: %d cells differ, first at %s: got %s want %s" % (
        what, len(bad), bad[0], got[tuple(bad[0])], ref[tuple(bad[0])])
