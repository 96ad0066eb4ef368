"""Closed-form pins of the oracle's infection probability (CPU).

The reference turns a susceptible replica's exposure E (seconds of contact
with infectious neighbours) into P = 1 - expf(logf(1 - p) * (E / 3600)),
with p the probability per hour of contact, whatever the step size
(tool/gpu_cpu_temporal_sim.cu:113-135 on the GPU, 439-458 on the CPU; the
commented-out line at cu:449 once divided by the step size instead).  So
whole hours of contact must give P = 1 - (1 - p)^h, and in general
P = 1 - (1 - p)^(E / 3600 h).  These tests check the oracle against those
powers, evaluated in FP64 (numpy's pow, no logarithm or exponential of the
oracle's), at step sizes other than one hour, and show that a build of the
oracle with the cu:449 divisor fails them.

Base of the power: the float 1 - p the reference computes (fl32(1 - p)),
since for p = 1e-4 the rounding of 1 - p alone moves P by about 6e-4
relative; the FP64 power of that float is the closed form.

Tolerance, per case: the fp32 exponent x = fl(fl(logf(a)) * fl(E / 3600))
carries at most 1 ulp (logf) + 0.5 (division) + 0.5 (product), under
2^-22 relative, which moves e^x by at most e^x |x| 2^-22; the exponential
rounded to float adds 0.5 ulp of e^x <= 2^-25 and 1 - y another 2^-25.  So
|P - P64| <= 2^-22 (|x| e^x + 1) covers every case.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PROBS = [1e-4, 0.03, 0.2, 0.5, 0.9]
HOURS = list(range(1, 41))
STEPS = [3600.0, 900.0, 1234.5, 86400.0]
# exposures that are not whole hours: one second, parts of hours, fractional
# sums of scaled weights, several days
ODD_E = [1.0, 59.0, 1800.0, 5400.5, 12345.678, 3 * 86400.0 + 17.0, 0.25, 7199.999]


def _closed_form(p32: np.float32, E: np.ndarray):
    a = np.float64(np.float32(1.0) - p32)
    h = np.asarray(E, np.float64) / 3600.0
    y = np.power(a, h)
    P = 1.0 - y
    x = h * np.log(a)  # for the tolerance only
    tol = 2.0 ** -22 * (np.abs(x) * y + 1.0)
    return P, tol


def _probability(oracle, library, p, step, E, c=None):
    E = np.asarray(E, np.float32)
    if c is None:
        c = np.zeros(E.size, np.int32)
    params = oracle.make_params(M=1, step_size=step, infect_prob=p)
    return oracle.infection_probability(params, c, E, library=library)


def _violations(oracle, library=None):
    """Cases where P departs from the closed form by more than the tolerance."""
    bad = []
    for step in STEPS:
        for p in PROBS:
            p32 = np.float32(p)
            for E in (np.array(HOURS, np.float64) * 3600.0, np.array(ODD_E, np.float64)):
                P = _probability(oracle, library, float(p32), step, E)
                ref, tol = _closed_form(p32, E.astype(np.float32))
                err = np.abs(P.astype(np.float64) - ref)
                for j in np.nonzero(~(err <= tol))[0]:
                    bad.append((step, p, float(E[j]), float(P[j]), float(ref[j])))
    return bad


def test_whole_hours_are_powers(oracle):
    """E = 3600 h seconds gives P = 1 - (1 - p)^h for h = 1..40 and p in
    {1e-4, 0.03, 0.2, 0.5, 0.9}, at step sizes of 1 h, 15 min, 1234.5 s and
    1 day; exposures that are not whole hours give 1 - (1 - p)^(E / 3600)."""
    bad = _violations(oracle)
    assert not bad, f"{len(bad)} cases off the closed form, e.g. {bad[:5]}"


def test_probability_does_not_depend_on_step(oracle):
    """The same exposure gives the same P, bit for bit, at every step size:
    the normalisation is per hour of contact (cu:127-128, 449-450)."""
    E = np.concatenate([np.array(HOURS, np.float32) * 3600.0, np.array(ODD_E, np.float32)])
    for p in PROBS:
        ref = _probability(oracle, None, p, 3600.0, E)
        for step in STEPS[1:]:
            np.testing.assert_array_equal(_probability(oracle, None, p, step, E), ref)


def test_zero_exposure_and_non_susceptible(oracle):
    """E = 0 gives P = 0 for p < 1; a replica whose countdown is not 0 gets
    P = 0 whatever its exposure (cu:444-455)."""
    for p in PROBS:
        for step in STEPS:
            assert _probability(oracle, None, p, step, [0.0]).tolist() == [0.0]
            P = _probability(oracle, None, p, step, [3600.0, 7200.0, 1.0],
                             c=np.array([1, 3600, 2 ** 30], np.int32))
            assert P.tolist() == [0.0, 0.0, 0.0]


def test_mutated_divisor_is_rejected(oracle, tmp_path):
    """A copy of the oracle built with the reference's commented-out divisor
    (E / step_size, cu:449) fails the closed-form checks at every step size
    other than one hour, and passes them at one hour: the checks pin the
    normalisation, not just the shape of the expression."""
    cc = shutil.which(os.environ.get("CC", "gcc"))
    if cc is None:
        pytest.skip("no C compiler")
    so = str(tmp_path / "liboracle_mut.so")
    src = os.path.join(ROOT, "oracle", "oracle.c")
    subprocess.run([cc, "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-DOR_HOUR_DIVISOR(p)=((p)->step_size)", src, "-o", so, "-lm"],
                   check=True, capture_output=True)
    mut = C.CDLL(so)
    P = C.POINTER
    mut.or_infection_probability.argtypes = [P(oracle.Params), C.c_int, P(C.c_int32),
                                             P(C.c_float)]
    mut.or_infection_probability.restype = None
    bad = _violations(oracle, library=mut)
    steps_caught = {b[0] for b in bad}
    assert steps_caught == set(STEPS) - {3600.0}, steps_caught
    # and the faithful build compiled the same way passes
    so2 = str(tmp_path / "liboracle_same.so")
    subprocess.run([cc, "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    src, "-o", so2, "-lm"], check=True, capture_output=True)
    same = C.CDLL(so2)
    same.or_infection_probability.argtypes = mut.or_infection_probability.argtypes
    same.or_infection_probability.restype = None
    assert not _violations(oracle, library=same)
