"""GPU parity of the uneven-length persistent batch (spx_batch_*).

Every channel's counts, statistics and peak bins must equal what
spx_fused_peaks gives for that channel alone (the same passes and reduction
order; record values to within a few ulps) and, on sampled channels, the
oracle's tie-tolerant peak list.  Graph replays are checked against the eager
result on every replay.
"""
import json
import os

import numpy as np
import pytest

from spectrum_amd import capi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

K = 64


def make_batch(ctx, n_chan, seed, lo=10, hi=17):
    L = capi.siggen_lengths(seed, n_chan, lo, hi)
    off, total, groups = capi.length_major_offsets(L)
    x = capi.siggen_uneven(capi.GEN_MULTI_TONE, seed, L, off, total)
    return L, off, total, groups, x, ctx.to_device(x)


def outputs(ctx, n):
    return ctx.alloc(n * K * 32), ctx.alloc(n * 8), ctx.alloc(n * 32), ctx.alloc(n * 16)


def read(ctx, bufs, n):
    pk, ct, gm, st = bufs
    return (pk.download(capi.MAXVALUE_DTYPE, (n, K)), ct.download(np.uint32, (n, 2)),
            gm.download(capi.MAXVALUE_DTYPE, (n,)), st.download(np.float32, (n, 4)))


def sequential(ctx, d_in, groups, n, refine=capi.REFINE_PARABOLIC):
    """Per-length spx_fused_peaks launches over the length-major layout -> arrays indexed by channel."""
    pk = np.zeros((n, K), capi.MAXVALUE_DTYPE)
    ct = np.zeros((n, 2), np.uint32)
    gm = np.zeros(n, capi.MAXVALUE_DTYPE)
    st = np.zeros((n, 4), np.float32)
    for L, (o, ids) in sorted(groups.items()):
        B = len(ids)
        bufs = outputs(ctx, B)
        ctx.fused_peaks(d_in.ptr + 8 * o, B, L, L, bufs[0], bufs[1], window=1, fs=1.0, nsigma=6.0, max_peaks=K,
                        gmax=bufs[2], stats=bufs[3], refine=refine)
        ctx.sync()
        p, c, g, s = read(ctx, bufs, B)
        pk[ids], ct[ids], gm[ids], st[ids] = p, c, g, s
    return pk, ct, gm, st


def assert_same(a, b, what, off_tol=1e-6):
    """Counts and peak bins identical; statistics and record values equal to within a few ulps (the LDS-tile
    body is the same source as the single-channel kernel, but instruction selection may fuse a
    multiply-add differently when it is inlined into the batch kernel)."""
    pa, ca, ga, sa = a
    pb, cb, gb, sb = b
    assert np.array_equal(ca, cb), f"{what}: counts differ"
    assert np.allclose(sa, sb, rtol=2e-6, atol=0), f"{what}: stats differ"
    assert np.array_equal(ga["index"], gb["index"]), f"{what}: global max bins differ"
    for c in range(len(ca)):
        n = ca[c, 0]
        ra, rb = pa[c, :n], pb[c, :n]
        if ra.tobytes() == rb.tobytes():
            continue
        assert np.array_equal(ra["index"], rb["index"]), f"{what}: channel {c} peak bins differ"
        scale = float(np.abs(rb["magnitude"]).max())
        for f in ("real", "imag", "magnitude"):
            assert np.max(np.abs(ra[f] - rb[f])) <= 4e-7 * scale, f"{what}: channel {c} field {f}"
        assert np.max(np.abs(ra["freq_offset"] - rb["freq_offset"])) <= off_tol, f"{what}: channel {c} freq_offset"
        dph = (ra["phase"] - rb["phase"] + 180.0) % 360.0 - 180.0
        assert np.max(np.abs(dph)) <= 1e-3, f"{what}: channel {c} phase"


def oracle_check(x, off, L, res, chans):
    pk, ct, gm, st = res
    for c in chans:
        n = int(L[c])
        X = orc.spectrum64(x[int(off[c]): int(off[c]) + n][None], n, 1)[0]
        recs, found, g, stats, mags = orc.fused_peaks(X, 1.0, nsigma=6.0, kcap=K)
        tau = 8 * 2.0 ** -24 * np.log2(n) * mags.max()
        ok, why = orc.compare_peak_lists(pk[c, : ct[c, 0]]["index"], recs["index"], mags, stats[2], tau, K,
                                         found_gpu=ct[c, 1], found_ref=found)
        assert ok, f"channel {c} (N={n}): {why}"
        assert abs(st[c, 2] - stats[2]) <= 1e-5 * stats[2] + tau


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_per_length_launches(ctx, graph):
    n = 96
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0103)
    print(f"[gpu] batch n={n} lengths={sorted(groups)} graph={graph}", flush=True)
    bufs = outputs(ctx, n)
    desc = ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3], max_peaks=K, graph=graph)
    plan = ctx.batch_plan(desc)
    info = plan.info()
    assert info["graph"] == graph and sum(b["channels"] for b in info["lengths"]) == n
    plan.execute()
    ctx.sync()
    got = read(ctx, bufs, n)
    ref = sequential(ctx, d_in, groups, n)
    assert_same(got, ref, "batch vs per-length")
    pick = [int(ids[0]) for _, (_, ids) in sorted(groups.items())]
    oracle_check(x, off, L, got, pick)
    plan.close()


@pytest.mark.parametrize("concurrent", [False, True])
def test_graph_replays_match_eager(ctx, tmp_path, concurrent):
    """Replays of the captured graph (with the LDS-tile launches on a forked
    and joined side stream when concurrent) equal the eager per-length result."""
    n = 40
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0203)
    bufs = outputs(ctx, n)
    ref = sequential(ctx, d_in, groups, n)
    cache = tmp_path / "conc.json"
    cache.write_text(json.dumps({"concurrent": int(concurrent),
                                 "lengths": [{"length": 1 << b, "tile": 4096} for b in range(10, 18)]}))
    eager = ctx.batch_plan(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3],
                                          max_peaks=K, graph=False, tune_cache=str(cache)))
    assert eager.info()["concurrent"] == concurrent
    eager.execute()
    ctx.sync()
    assert_same(read(ctx, bufs, n), ref, "eager")
    plan = ctx.batch_plan(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3],
                                         max_peaks=K, graph=True, tune_cache=str(cache)))
    for rep in range(3):
        # clobber the outputs with eager work between replays
        ctx.fused_peaks(d_in, 1, 1024, 1024, bufs[0], bufs[1], max_peaks=K, gmax=bufs[2], stats=bufs[3])
        ctx.sync()
        plan.execute()
        ctx.sync()
        assert_same(read(ctx, bufs, n), ref, f"replay {rep}")
    plan.close()
    eager.close()


def test_alternate_radix_order_and_autotune(ctx, tmp_path):
    n = 64
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0303)
    bufs = outputs(ctx, n)
    # a cache that forces the leftover-radix-first order everywhere
    cache = tmp_path / "order1.json"
    cache.write_text(json.dumps({"device": "any", "lengths": [{"length": 1 << b, "order": 1} for b in range(10, 18)]}))
    desc = ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3], max_peaks=K,
                          tune_cache=str(cache))
    plan = ctx.batch_plan(desc)
    orders = {b["length"]: b["order"] for b in plan.info()["lengths"]}
    for Ln, o in orders.items():
        if Ln >= 32768:
            assert o == 0, orders  # the split path of the long lengths has one radix order
        else:
            assert o == (0 if (Ln.bit_length() - 1) % 4 == 0 else 1), orders
    plan.execute()
    ctx.sync()
    got = read(ctx, bufs, n)
    pick = [int(ids[i]) for _, (_, ids) in sorted(groups.items()) for i in range(min(2, len(ids)))]
    oracle_check(x, off, L, got, pick)
    plan.close()
    # the autotuner measures every (order, tile shape) of the LDS-tile lengths
    # and 1 / 2 / 4 launch groups of the split lengths, and writes its cache
    tuned = tmp_path / "tuned.json"
    ctx.batch_autotune(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], max_peaks=K), str(tuned))
    data = json.loads(tuned.read_text())
    assert {e["length"] for e in data["lengths"]} == set(int(v) for v in np.unique(L))
    sets = data["tile_sets"]
    if "whole_us" in sets[0]:
        # with split lengths present: tile set and concurrency from whole-batch
        # times; the default (first set, in sequence) unless another is 2 % faster
        t = [v for s_ in sets for v in (s_["whole_us"], s_["whole_concurrent_us"])]
        k = min(range(len(t)), key=lambda i: (t[i], i))
        k = k if t[k] < 0.98 * t[0] else 0
        win = sets[k // 2]
        assert data["concurrent"] == k % 2
    else:
        win = min(sets, key=lambda t: t["us"])
    for e in data["lengths"]:
        cands = e["candidates"]
        assert all(c["us"] > 0 for c in cands)
        chosen = [c for c in cands if (c["order"], c["tile"], c["groups"]) == (e["order"], e["tile"], e["groups"])]
        assert len(chosen) == 1 and chosen[0]["us"] == e["us"], (e, cands)
        Ln = e["length"]
        if Ln >= 32768:
            # launch groups of a split length: the fastest alone, one group
            # unless another count is 2 % faster
            assert {c["groups"] for c in cands} <= {1, 2, 4} and all(c["order"] == 0 for c in cands)
            fastest = min(c["us"] for c in cands)
            assert e["us"] == (fastest if fastest < 0.98 * cands[0]["us"] else cands[0]["us"])
        else:
            two = (Ln.bit_length() - 1) % 4 != 0
            want = {(o, t) for t in (16384, 8192, 4096) if t >= Ln for o in ((0, 1) if two else (0,))}
            assert {(c["order"], c["tile"]) for c in cands} == want, (Ln, cands)
            # tile shapes: those of the fastest set timed over all LDS-tile
            # lengths together; order 1 on that shape only where 2 % faster
            assert e["tile"] == win["tiles"][str(Ln)]
            t = {c["order"]: c["us"] for c in cands if c["tile"] == e["tile"]}
            assert e["order"] == (1 if 1 in t and t[1] < 0.98 * t[0] else 0)
    plan = ctx.batch_plan(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], max_peaks=K, tune_cache=str(tuned)))
    info = plan.info()
    for b in info["lengths"]:
        e = [x for x in data["lengths"] if x["length"] == b["length"]][0]
        assert b["order"] == e["order"]
        if b["path"] == "lds-tile":
            assert b["tile"] == e["tile"] and b["channels_per_tile"] == e["tile"] // b["length"]
    plan.close()


@pytest.mark.parametrize("tile", [16384, 8192, 4096])
def test_tile_shapes_equal_per_length_launches(ctx, tmp_path, tile):
    """Every tile shape (and both radix orders where a length has two), with the
    split lengths in two launch groups, gives the per-length launches' results."""
    n = 96
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0403)
    ref = sequential(ctx, d_in, groups, n)
    for order, conc in ((0, 0), (1, 1)):
        print(f"[gpu] batch tile={tile} order={order} concurrent={conc} lengths={sorted(groups)}", flush=True)
        cache = tmp_path / f"t{tile}_{order}.json"
        cache.write_text(json.dumps({"concurrent": conc,
                                     "lengths": [{"length": 1 << b, "order": order, "tile": tile, "groups": 2}
                                                 for b in range(10, 18)]}))
        bufs = outputs(ctx, n)
        plan = ctx.batch_plan(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3],
                                             max_peaks=K, tune_cache=str(cache)))
        info = plan.info()
        assert info["concurrent"] == bool(conc)
        for b in info["lengths"]:
            if b["path"] == "lds-tile":
                assert b["tile"] == (tile if tile >= b["length"] else 16384), info
            else:
                assert b["groups"] == 2, info
        assert {t["tile"] for t in info["launches"]} == {b["tile"] for b in info["lengths"] if b["path"] == "lds-tile"}
        plan.execute()
        ctx.sync()
        assert_same(read(ctx, bufs, n), ref, f"tile {tile} order {order}")
        plan.close()


def test_full_uneven_config_properties(ctx):
    """4096 channels, lengths log-uniform 2^10..2^17 (about 1.07 GB): equal to per-length launches."""
    n = 4096
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0003)
    print(f"[gpu] uneven config: {n} channels, {total * 8 / 1e9:.2f} GB", flush=True)
    bufs = outputs(ctx, n)
    plan = ctx.batch_plan(ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3], max_peaks=K))
    plan.execute()
    ctx.sync()
    got = read(ctx, bufs, n)
    ref = sequential(ctx, d_in, groups, n)
    assert_same(got, ref, "uneven config")
    pk, ct, gm, st = got
    assert np.all(ct[:, 0] == np.minimum(ct[:, 1], K))
    for c in range(0, n, 97):
        idx = pk[c, : ct[c, 0]]["index"]
        assert np.all(np.diff(idx.astype(np.int64)) > 0)
        assert np.all(pk[c, : ct[c, 0]]["magnitude"] > st[c, 2])
    oracle_check(x, off, L, got, [int(ids[-1]) for Lx, (o, ids) in sorted(groups.items())])
    plan.close()


@pytest.mark.parametrize("concurrent", [False, True])
def test_learned_refinement_equals_per_length_launches(ctx, tmp_path, concurrent):
    """refine = LEARNED in a batch plan: the LDS-tile channels' features feed one
    refiner launch over those channels (channel indirection, per-channel bin
    width), the split channels are refined per launch group; the records equal
    per-length spx_fused_peaks calls with learned refinement."""
    from spectrum_amd.refiner import DEFAULT_WEIGHTS, Refiner

    n = 96
    L, off, total, groups, x, d_in = make_batch(ctx, n, 0x5EED0503)
    bufs = outputs(ctx, n)
    desc = ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3], max_peaks=K,
                          refine=capi.REFINE_LEARNED)
    ctx.set_refiner(None)
    with pytest.raises(ValueError):
        ctx.batch_plan(desc)  # no refiner set
    r = Refiner(ctx).load(DEFAULT_WEIGHTS)
    r.export()
    ctx.set_refiner(r.deployed)
    try:
        ref = sequential(ctx, d_in, groups, n, refine=capi.REFINE_LEARNED)
        par = sequential(ctx, d_in, groups, n)
        cache = tmp_path / "c.json"
        cache.write_text(json.dumps({"concurrent": int(concurrent), "lengths": []}))
        desc = ctx.batch_desc(d_in, off, L, bufs[0], bufs[1], gmax=bufs[2], stats=bufs[3], max_peaks=K,
                              refine=capi.REFINE_LEARNED, tune_cache=str(cache))
        print(f"[gpu] learned batch n={n} lengths={sorted(groups)} concurrent={concurrent}", flush=True)
        plan = ctx.batch_plan(desc)
        assert plan.info()["concurrent"] == concurrent
        for rep in range(2):
            plan.execute()
            ctx.sync()
            got = read(ctx, bufs, n)
            # a feature rounded to fp16 from a spectrum value one ulp apart can
            # move a learned offset by ~1e-4 bin
            assert_same(got, ref, f"learned batch replay {rep}", off_tol=1e-3)
        plan.close()
        # the refiner did act: offsets differ from the parabolic ones somewhere
        pk, ct = got[0], got[1]
        diff = sum(int(np.any(pk[c, : ct[c, 0]]["freq_offset"] != par[0][c, : ct[c, 0]]["freq_offset"]))
                   for c in range(n))
        assert diff > 0
    finally:
        ctx.set_refiner(None)


def test_batch_rejects_bad_lengths(ctx):
    d = ctx.alloc(4096 * 8)
    bufs = outputs(ctx, 2)
    for bad in ([1024, 512], [1024, 3000], [1024, 1 << 18]):
        desc = ctx.batch_desc(d, np.array([0, 1024], np.uint64), np.array(bad, np.uint32), bufs[0], bufs[1])
        with pytest.raises(ValueError):
            ctx.batch_plan(desc)
    desc = ctx.batch_desc(d, np.array([0, 1024], np.uint64), np.array([1024, 1024], np.uint32), bufs[0], bufs[1],
                          max_peaks=0)
    with pytest.raises(ValueError):
        ctx.batch_plan(desc)
