"""The coefficient table of the batched form's matrix functions
(letkf_rational_table_entry): Zolotarev's rational approximation
x^-1/2 ~ d + sum_l a_l / (x + s_l) on [1, 2^(cls+1)].

Pinned by the closed form x^-1/2: on a dense logarithmic grid of every
class's interval the relative error stays below the table's targets
(1e-15 for the nmem x nmem form, 1e-16 for the nobsl x nobsl form, 1e-10
and 1e-11 for the same forms after an FP32 reduction) up to the rounding
of a double evaluation, and, in 40-digit arithmetic (mpmath)
on the stored double coefficients, below the targets themselves for the
classes the workloads reach.  The residues and poles are positive (a
Stieltjes function's best approximation), and the order grows with the
interval, as the theory has it (~log(kappa) poles).
"""
import ctypes as C

import numpy as np
import pytest

from sbpom_letkf_v2_amd import _lib

CLASSES = 24
TARGET = {0: 1e-15, 1: 1e-16, 2: 1e-10, 3: 1e-11}


def _entry(kind, cls):
    L = _lib.lib()
    d = C.c_double()
    a = (C.c_double * 40)()
    s = (C.c_double * 40)()
    n = L.letkf_rational_table_entry(kind, cls, C.byref(d), a, s)
    assert n >= 0
    return n, d.value, np.array(a[:n]), np.array(s[:n])


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_table_meets_its_targets_in_double(kind):
    prev = 0
    for cls in range(CLASSES):
        n, d, a, s = _entry(kind, cls)
        if n == 0:
            # a class without an entry: every larger one is missing too
            assert all(_entry(kind, c)[0] == 0 for c in range(cls, CLASSES))
            assert cls >= 20
            break
        assert n >= prev and (a > 0).all() and (s > 0).all() and d > 0
        prev = n
        kap = 2.0 ** (cls + 1)
        x = np.exp(np.linspace(0.0, np.log(kap), 40001))
        r = d + (a[None, :] / (x[:, None] + s[None, :])).sum(1)
        err = np.abs(r * np.sqrt(x) - 1.0).max()
        assert err <= max(2e-15, 1.1 * TARGET[kind]), (kind, cls, n, err)


@pytest.mark.parametrize("kind,cls", [(0, 0), (0, 5), (0, 9), (1, 0), (1, 3), (1, 6), (2, 5), (3, 6)])
def test_table_error_in_high_precision(kind, cls):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    n, d, a, s = _entry(kind, cls)
    A = [mp.mpf(float(v)) for v in a]
    S = [mp.mpf(float(v)) for v in s]
    D = mp.mpf(d)
    kap = mp.mpf(2) ** (cls + 1)
    target = TARGET[kind]
    worst = mp.mpf(0)
    for q in range(1201):
        x = mp.e ** (mp.log(kap) * q / 1200)
        r = D + mp.fsum(ai / (x + si) for ai, si in zip(A, S))
        worst = max(worst, abs(r * mp.sqrt(x) - 1))
    # the stored doubles add at most a few units of 1e-17 to the target
    assert float(worst) <= target * 1.1, (kind, cls, float(worst))


def test_looser_kinds_need_fewer_poles():
    # the FP32-reduction kinds: about a third fewer poles per class
    for kind, loose in ((0, 2), (1, 3)):
        for cls in (0, 3, 6, 9):
            n, nl = _entry(kind, cls)[0], _entry(loose, cls)[0]
            assert 0 < nl < n, (kind, cls, n, nl)


def test_table_entry_rejects_bad_arguments():
    L = _lib.lib()
    d = C.c_double()
    buf = (C.c_double * 40)()
    assert L.letkf_rational_table_entry(4, 0, C.byref(d), buf, buf) < 0
    assert L.letkf_rational_table_entry(0, CLASSES, C.byref(d), buf, buf) < 0
    assert L.letkf_rational_table_entry(0, 0, None, buf, buf) < 0
