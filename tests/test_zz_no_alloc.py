"""spx_fused_peaks_noalloc (fixed device workspace, no allocation) against
spx_fused_peaks: the same records, counts, global maxima and statistics,
byte for byte, on two calls in a row (the second with a smaller input), on
the long path (several launch groups of the fixed workspace) and the short
path; the call's error paths (tables not prepared, learned features larger
than the workspace) allocate nothing and leave the context usable."""
import numpy as np
import pytest

from spectrum_amd import capi

pytestmark = pytest.mark.gpu


def _outputs(ctx, B, K):
    return ctx.alloc(B * K * 32), ctx.alloc(B * 8), ctx.alloc(B * 32), ctx.alloc(B * 16)


def _call(ctx, d_in, B, N, K, bufs, noalloc, **kw):
    pk, ct, g, st = bufs
    # poison the outputs so that a record the call does not write shows up
    pk.upload(np.full(B * K * 32, 0xA5, np.uint8))
    ctx.fused_peaks(d_in, B, N, N, pk, ct, max_peaks=K, gmax=g, stats=st, noalloc=noalloc, **kw)
    ctx.sync()
    return (pk.download(np.uint8, (B, K * 32)), ct.download(np.uint32, (B, 2)), g.download(np.uint8, (B, 32)),
            st.download(np.float32, (B, 4)))


def _same(a, b):
    pa, ca, ga, sa = a
    pb, cb, gb, sb = b
    assert np.array_equal(ca, cb)
    for c in range(ca.shape[0]):
        n = int(ca[c, 0]) * 32
        assert pa[c, :n].tobytes() == pb[c, :n].tobytes(), c
    assert ga.tobytes() == gb.tobytes() and sa.tobytes() == sb.tobytes()


def test_noalloc_equals_fused_peaks_two_calls(ctx):
    # call 1: 1024 x 65536 (21 launch groups of the fixed workspace)
    B, N, K = 1024, 65536, 64
    print(f"[case] noalloc first call {B}x{N}", flush=True)
    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0202, B, N)
    d_in = ctx.to_device(x)
    bufs = _outputs(ctx, B, K)
    ref = _call(ctx, d_in, B, N, K, bufs, False)  # also creates this length's tables
    got = _call(ctx, d_in, B, N, K, bufs, True)
    _same(got, ref)
    assert np.all(got[1][:, 1] >= 1)
    # call 2, smaller input: 37 x 4096 on the short path
    B2, N2 = 37, 4096
    print(f"[case] noalloc second call {B2}x{N2}", flush=True)
    x2 = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0203, B2, N2)
    d2 = ctx.to_device(x2)
    bufs2 = _outputs(ctx, B2, 32)
    ref2 = _call(ctx, d2, B2, N2, 32, bufs2, False)
    got2 = _call(ctx, d2, B2, N2, 32, bufs2, True)
    _same(got2, ref2)
    # and a smaller long-path call after that: 5 x 262144, band search
    print("[case] noalloc third call 5x262144", flush=True)
    x3 = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0204, 5, 1 << 18)
    d3 = ctx.to_device(x3)
    bufs3 = _outputs(ctx, 5, 16)
    ref3 = _call(ctx, d3, 5, 1 << 18, 16, bufs3, False, s0=100, s1=150000)
    got3 = _call(ctx, d3, 5, 1 << 18, 16, bufs3, True, s0=100, s1=150000)
    _same(got3, ref3)


def test_noalloc_errors_allocate_nothing(ctx):
    print("[case] noalloc errors", flush=True)
    # a window over an n_point no earlier call used: its table does not exist
    x = capi.siggen(capi.GEN_THREE_TONES, 0x5EED0205, 2, 3001)
    d_in = ctx.to_device(x)
    pk, ct = ctx.alloc(2 * 8 * 32), ctx.alloc(2 * 8)
    with pytest.raises(ValueError, match="not in the context"):
        ctx.fused_peaks(d_in, 2, 3001, 4096, pk, ct, max_peaks=8, window=3, noalloc=True)
    # the context still works, and after an allocating call the same no-alloc call runs
    ctx.fused_peaks(d_in, 2, 3001, 4096, pk, ct, max_peaks=8, window=3)
    ctx.fused_peaks(d_in, 2, 3001, 4096, pk, ct, max_peaks=8, window=3, noalloc=True)
    ctx.sync()
    assert np.all(ct.download(np.uint32, (2, 2))[:, 1] >= 1)


def test_noalloc_learned_refinement(ctx):
    """Learned refinement keeps its peak features in the fixed workspace:
    equal to spx_fused_peaks within it, SPX_ECAP beyond it."""
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    print("[case] noalloc learned", flush=True)
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    ctx.set_refiner(r.deployed)
    try:
        B, N, K = 512, 8192, 64
        x = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0206, B, N)
        d_in = ctx.to_device(x)
        bufs = _outputs(ctx, B, K)
        ref = _call(ctx, d_in, B, N, K, bufs, False, refine=capi.REFINE_LEARNED)
        got = _call(ctx, d_in, B, N, K, bufs, True, refine=capi.REFINE_LEARNED)
        _same(got, ref)
        # 4096 channels x 1024 records x 32 bytes of features = 128 MiB > 64 MiB
        B2, N2, K2 = 4096, 1024, 1024
        x2 = capi.siggen(capi.GEN_FRACTIONAL, 0x5EED0207, B2, N2)
        d2 = ctx.to_device(x2)
        pk, ct = ctx.alloc(B2 * K2 * 32), ctx.alloc(B2 * 8)
        ctx.fused_peaks(d2, B2, N2, N2, pk, ct, max_peaks=K2)  # creates the tables
        with pytest.raises(capi.SpxError, match="fixed workspace"):
            ctx.fused_peaks(d2, B2, N2, N2, pk, ct, max_peaks=K2, refine=capi.REFINE_LEARNED, noalloc=True)
    finally:
        ctx.set_refiner(None)
