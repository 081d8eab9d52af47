"""Host-side plan and Stockham twiddle tables of the 2^19 .. 2^24 path
(spx_large_plan / spx_large_tables in libspectrum_mi355x.so, no GPU needed):
every radix plan multiplies out to N with LDS passes the kernels exist for,
the tables hold W_m^e = exp(-2 pi i e / m) rounded once from FP64 in the
layout the LDS-pass kernel reads ([jj][k], then the powers 1, 2, 3, 4, 8, 12
as [i][k]), and the kernel's factorisation W_(NS L)^(e k) = W_(NS L)^(jj k)
U^r (e = jj + r L/16, U = W_(16 NS)^k, U^r with at most one product)
reproduces the closed form."""
import ctypes as C

import numpy as np
import pytest

from spectrum_amd import capi


@pytest.fixture(scope="module")
def lib():
    lib = C.CDLL(capi.LIB_PATH)
    lib.spx_large_tables.restype = C.c_uint64
    lib.spx_large_tables.argtypes = [C.c_uint32, C.c_void_p]
    lib.spx_large_plan.argtypes = [C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return lib


def plan(lib, n):
    r0, np_ = C.c_int(), C.c_int()
    lgs = (C.c_int * 3)()
    rc = lib.spx_large_plan(n, C.byref(r0), lgs, C.byref(np_))
    return rc, r0.value, [lgs[i] for i in range(np_.value)]


@pytest.mark.parametrize("lg", range(19, 25))
def test_plan_and_tables(lib, lg):
    N = 1 << lg
    rc, r0, lgs = plan(lib, N)
    assert rc == 0 and r0 in (3, 4) and r0 + sum(lgs) == lg
    assert all(x in (8, 9, 10) for x in lgs)
    nf = lib.spx_large_tables(N, None)
    t = np.zeros(nf, np.float32)
    assert lib.spx_large_tables(N, t.ctypes.data) == nf
    tw = t[0::2].astype(np.float64) + 1j * t[1::2].astype(np.float64)
    rng = np.random.default_rng(lg)
    off, ns = 0, 1 << r0
    for l in lgs:
        L = 1 << l
        A = tw[off: off + ns * (L // 16)]
        off += ns * (L // 16)
        B = tw[off: off + 6 * ns]
        off += 6 * ns
        # closed form at random entries
        jj = rng.integers(0, L // 16, 200)
        k = rng.integers(0, ns, 200)
        m = ns * L
        assert np.max(np.abs(A[jj * ns + k] - np.exp(-2j * np.pi * ((jj * k) % m) / m))) < 1.2e-7
        pw = np.array([1, 2, 3, 4, 8, 12])
        i = rng.integers(0, 6, 200)
        assert np.max(np.abs(B[i * ns + k] - np.exp(-2j * np.pi * ((pw[i] * k) % (16 * ns)) / (16 * ns)))) < 1.2e-7
        # the kernel's product form of the Stockham twiddle
        e = rng.integers(0, L, 200)
        jj, r = e % (L // 16), e // (L // 16)
        lo, hi = r & 3, r >> 2
        u = [B[q * ns + k] for q in range(6)]
        ur = np.where(r == 0, 1.0, 0j)
        for q in range(200):
            if r[q] == 0:
                continue
            a = u[lo[q] - 1][q] if lo[q] else 1.0
            b = u[2 + hi[q]][q] if hi[q] else 1.0
            ur[q] = a * b
        w = A[jj * ns + k] * ur
        assert np.max(np.abs(w - np.exp(-2j * np.pi * ((e * k) % m) / m))) < 4e-7
        ns *= L
    assert off == len(tw)


def test_plan_rejects_other_lengths(lib):
    for n in (1 << 18, (1 << 19) + 1, 1 << 25, 3 << 19):
        assert plan(lib, n)[0] != 0
    assert lib.spx_large_tables(1 << 18, None) == 0
