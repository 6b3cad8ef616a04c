"""The reference's Eq. 9 known-answer cases (tests/golden/relocation_cases.json, from the
reference's tests/test_relocation.cpp) on the relocation arithmetic of the CPU oracle; the
GPU relocation kernel is checked against the same cases in tests/test_gpu_mcmc.py."""
import ctypes as C
import json
import os

import numpy as np

import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "relocation_cases.json")))


def check_eq9(eq9):
    """eq9(o_old, n) -> (o_new, s_new) with s_old = 1; asserts the reference's invariants."""
    lo, hi = CASES["integral_range"]
    x = np.arange(lo, hi, CASES["integral_dx"])
    for o in CASES["opacities"]:
        for n in CASES["ensemble_sizes"]:
            on, sn = eq9(np.float32(o), n)
            centre = 1.0 - (1.0 - float(on)) ** n
            assert abs(centre - float(np.float32(o))) <= CASES["centre_alpha_tol"], (o, n)
            a = float(on) * np.exp(-x * x / (2.0 * float(sn) ** 2))
            integral = ((1.0 - (1.0 - a) ** n) * CASES["integral_dx"]).sum()
            want = float(np.float32(o)) * np.sqrt(2 * np.pi)
            assert abs(integral / want - 1.0) <= CASES["integral_rel_tol"], (o, n, integral / want)
            if n == 1:
                assert abs(sn - 1.0) <= CASES["noop_tol"] and abs(on - np.float32(o)) <= CASES["noop_tol"]
            else:
                assert sn > np.sqrt(1.0 / n), (o, n)


def test_oracle_eq9_known_answers():
    L = O.lib()

    def eq9(o, n):
        s_old = np.ones(3, np.float32)
        s_new = np.zeros(3, np.float32)
        on = C.c_float()
        L.oracle_eq9(float(o), s_old.ctypes.data, n, C.byref(on), s_new.ctypes.data)
        assert s_new[0] == s_new[1] == s_new[2]
        return on.value, float(s_new[0])
    check_eq9(eq9)
