"""The infection draw's fast decision (CPU).

The kernel compares a uniform u with an estimate of the reference's
P = 1 - expf(log(1-p) * (E / 3600)) (cu:113-135) and evaluates the reference
expression only when |P_est - u| <= 1e-5.  That is exact as long as the
estimate is within the band of the reference's P.  The estimate is
1 - 2^(fl(log(1-p) * fl(E * fl(1/3600))) * log2 e) from the hardware base-2
exponential, which is within 1 ulp; here that exponential is the float32
exp2 of numpy moved by -1, 0 and +1 ulp, and P comes from the oracle's
restatement of the reference expression.  The measured gap must stay below
5e-7, the bound the kernel's comment derives, 20 times inside the band."""
import ctypes as C

import numpy as np

BAND = 1.0e-5
BOUND = 5.0e-7


def _logf():
    libm = C.CDLL("libm.so.6")
    libm.logf.argtypes = [C.c_float]
    libm.logf.restype = C.c_float
    return libm.logf


def test_estimate_within_bound(oracle):
    f32 = np.float32
    logf = _logf()
    rng = np.random.default_rng(0)
    probs = np.concatenate([[1e-6, 1e-4, 0.001, 0.01, 0.03, 0.05, 0.2, 0.5, 0.9, 0.999, 0.999999],
                            rng.uniform(0.0, 1.0, 12)]).astype(np.float32)
    E = np.concatenate([np.arange(1, 400, dtype=np.float64) * 600.0,          # whole windows
                        rng.integers(1, 1 << 24, 600).astype(np.float64),    # integer seconds
                        rng.uniform(0, 2e5, 600)]).astype(np.float32)         # float-mode sums
    inv3600 = f32(1.0) / f32(3600.0)
    log2e = f32(1.4426950408889634)
    worst = 0.0
    for p in probs:
        lb = f32(logf(f32(1.0) - p))
        x_ref = (lb * (E / f32(3600.0))).astype(np.float32)
        P_ref = np.array([f32(1.0) - f32(oracle.expf(float(x))) for x in x_ref], np.float32)
        x_est = (lb * (E * inv3600)).astype(np.float32)
        e = np.exp2((x_est * log2e).astype(np.float32)).astype(np.float32)
        for ex in (np.nextafter(e, f32(0.0)), e, np.nextafter(e, f32(2.0))):
            P_est = (f32(1.0) - ex).astype(np.float32)
            gap = np.abs(P_est.astype(np.float64) - P_ref.astype(np.float64))
            worst = max(worst, float(gap.max()))
    assert worst < BOUND < BAND, worst


def test_band_decides_like_the_reference(oracle):
    """Outside the band the estimate's comparison equals the reference's:
    uniforms placed just outside and just inside the band around P."""
    f32 = np.float32
    logf = _logf()
    lb = f32(logf(f32(1.0) - f32(0.03)))
    E = (np.arange(1, 200, dtype=np.float32) * f32(600.0)).astype(np.float32)
    P_ref = np.array([f32(1.0) - f32(oracle.expf(float(f32(lb * f32(x / f32(3600.0))))))
                      for x in E], np.float32)
    x_est = (lb * (E * (f32(1.0) / f32(3600.0)))).astype(np.float32)
    P_est = (f32(1.0) - np.exp2((x_est * f32(1.4426950408889634)).astype(np.float32))).astype(np.float32)
    for off in (1.01 * BAND, -1.01 * BAND, 0.5 * BAND, -0.5 * BAND, 0.0):
        u = (P_est + f32(off)).astype(np.float32)
        outside = np.abs(P_est - u) > f32(BAND)
        decided = np.where(outside, P_est > u, P_ref > u)
        np.testing.assert_array_equal(decided, P_ref > u)
