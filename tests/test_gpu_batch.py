"""Batched ragged solve: every pair of a mixed batch equals its single-pair solve."""
import sys

import numpy as np
import pytest

from pdflow import Batch, PDFlowError, PinnedBuffer, Solver, synth_pair_float

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(rows, seed, iters=None):
    cam = 480 // rows
    I1, Z1, I2, Z2 = synth_pair_float(seed, 480 // cam, 640 // cam)
    g = Solver(rows, cam)
    return (I1, Z1, I2, Z2), g.solve_pair(I1, Z1, I2, Z2, iters=iters)


@pytest.mark.parametrize("k,graph,shape", [(4, True, 0), (2, False, 0), (0, True, 0), (4, True, 1), (3, False, 1),
                                           (5, True, 1), (6, False, 1)])
def test_mixed_batch_matches_single_pairs(gpu_available, k, graph, shape):
    rows = [480, 240, 120, 60, 240, 30, 120, 15, 480]
    say(f"mixed batch of {len(rows)} pairs, k={k}, graph={graph}, shape={shape}")
    b = Batch(rows)
    b.set_tb_shape(shape)
    if k:
        b.set_iteration_kernel(k, 4800)
    else:
        b.set_iteration_kernel(0, 0, level_kernel=False)
    want = []
    for i, r in enumerate(rows):
        inputs, flow = single(r, 500 + i)
        b.set_inputs(i, *inputs)
        want.append(flow)
    for rep in range(2):
        b.solve(graph=graph)
        b.synchronize()
        for i, r in enumerate(rows):
            for a, w in zip(b.flow(i), want[i]):
                np.testing.assert_array_equal(a, w, err_msg=f"pair {i} rows {r} rep {rep}")
    say(f"  {b.launch_count()} launches per batched solve")


def test_batch_pyramids_match_the_oracle(gpu_available):
    """Every pyramid plane a batched solve builds (both frames; colour, depth, xx, yy; every
    level; pairs of four sizes) equals the oracle's bit for bit, and level 0 of a pair's colour
    and depth pyramids is its input planes themselves."""
    from oracle import Oracle
    from pdflow import lib as plib
    rows = [480, 240, 120, 60]
    say(f"batched pyramids of {rows} against the oracle")
    b = Batch(rows, iters=[[1] * 6, [1] * 5, [1] * 4, [1] * 3])
    oracles = []
    for i, r in enumerate(rows):
        cam = 480 // r
        I1, Z1, I2, Z2 = synth_pair_float(900 + i, 480 // cam, 640 // cam)
        b.set_inputs(i, I1, Z1, I2, Z2)
        o = Oracle(r, cam, threads=8)
        o.copy_new_frames(I1, Z1)
        o.gaussian_pyramid()
        o.copy_new_frames(I2, Z2)
        o.gaussian_pyramid()
        oracles.append(o)
    b.solve(graph=True)
    b.synchronize()
    for i, o in enumerate(oracles):
        I1d, Z1d, I2d, Z2d = b.inputs_device(i)
        for kind, addr in (("colour_old", I1d), ("depth_old", Z1d), ("colour", I2d), ("depth", Z2d)):
            ptr = plib.ctypes.c_void_p()
            r_, c_ = plib.ctypes.c_uint(), plib.ctypes.c_uint()
            assert b.lib.pdf_batch_pyramid_device(b.ctx, i, plib.PYR_KINDS[kind], 0, plib.ctypes.byref(ptr),
                                                  plib.ctypes.byref(r_), plib.ctypes.byref(c_)) == 0
            assert ptr.value == addr, f"pair {i}: level-0 {kind} is not the input plane"
        for lvl in range(o.pyr_levels):
            for kind in plib.PYR_KINDS:
                got = b.pyramid(i, kind, lvl)[0]
                np.testing.assert_array_equal(got, o.get_pyramid(kind, lvl), err_msg=f"pair {i} {kind}[{lvl}]")


def test_continue_equals_a_longer_finest_level(gpu_available):
    """pdf_batch_continue: n more finest-level iterations from the solved state give exactly
    the flow of a solve whose finest level ran n more; continuing twice adds up; the next
    solve starts afresh."""
    rows = [240, 120, 240, 60]
    scheds = [[10, 11, 13, 17, 21], [3, 5, 7, 9], [40, 55, 70, 85, 100], [6, 9, 12]]
    cont = [0, 2, 3]                      # pair 1 is not continued
    say("continuation of the finest level: +7 then +5 against +12 in one solve")
    a = Batch(rows, iters=scheds)
    longer = [list(sc) for sc in scheds]
    for i in cont:
        longer[i][-1] += 12
    b = Batch(rows, iters=longer)
    for i, r in enumerate(rows):
        inputs = single(r, 900 + i)[0]
        a.set_inputs(i, *inputs)
        b.set_inputs(i, *inputs)
    a.solve(graph=True)
    a.synchronize()
    first = [a.flow(i) for i in range(len(rows))]
    a.continue_finest(cont, 7)
    a.continue_finest(cont, 5)
    a.synchronize()
    b.solve(graph=True)
    b.synchronize()
    for i, r in enumerate(rows):
        for x, y in zip(a.flow(i), b.flow(i)):
            np.testing.assert_array_equal(x, y, err_msg=f"pair {i} rows {r}")
    ea, ra = a.metrics()
    eb, rb = b.metrics()
    np.testing.assert_array_equal(ea, eb)
    np.testing.assert_array_equal(ra, rb)
    # the metrics of a subset, in the listed order, are the full batch's for those pairs
    es, rs = a.metrics_pairs([3, 0])
    np.testing.assert_array_equal(es, ea[[3, 0]])
    np.testing.assert_array_equal(rs, ra[[3, 0]])
    # a new solve of batch a replays its own schedule from scratch
    a.solve(graph=True)
    a.synchronize()
    for i in range(len(rows)):
        for x, y in zip(a.flow(i), first[i]):
            np.testing.assert_array_equal(x, y)
    with pytest.raises(Exception):
        a.continue_finest([0, 0], 3)


def test_batch_with_custom_schedules(gpu_available):
    """Per-pair iteration counts (as the learned warm start uses), odd counts included."""
    rows = [240, 120, 240]
    scheds = [[10, 11, 13, 17, 21], [3, 5, 7, 9], [40, 55, 70, 85, 100]]
    say("batch with custom schedules")
    b = Batch(rows, iters=scheds)
    for i, r in enumerate(rows):
        inputs, flow = single(r, 700 + i, iters=scheds[i])
        b.set_inputs(i, *inputs)
        b.solve()
        b.synchronize()
        for a, w in zip(b.flow(i), flow):
            np.testing.assert_array_equal(a, w)


def test_autotune_picks_a_config_and_caches_it(gpu_available, tmp_path):
    """The autotuner times every region shape and K on the batch itself, keeps the fastest,
    and a second batch of the same composition takes the cached choice; the flow is the same
    for every choice."""
    rows = [480, 240, 120]
    say("autotune on a 3-pair batch")
    cache = tmp_path / "tune.txt"
    b = Batch(rows)
    want = []
    for i, r in enumerate(rows):
        inputs, flow = single(r, 900 + i)
        b.set_inputs(i, *inputs)
        want.append(flow)
    shape, k, ms = b.autotune(cache)
    tile = b.setup_tile()
    assert shape in (0, 1) and k in (3, 4, 5, 6) and ms > 0 and 0 <= tile < len(Batch.SETUP_TILES)
    b.solve(graph=True)
    b.synchronize()
    for i in range(len(rows)):
        for a, w in zip(b.flow(i), want[i]):
            np.testing.assert_array_equal(a, w)
    lines = cache.read_text().splitlines()
    fields = lines[0].split("\t")[1].split()
    assert len(lines) == 1 and fields[:2] == [str(shape), str(k)] and fields[3] == str(tile)
    b2 = Batch(rows)
    for i, r in enumerate(rows):
        b2.set_inputs(i, *single(r, 900 + i)[0])
    assert b2.autotune(cache)[:2] == (shape, k) and b2.setup_tile() == tile
    assert len(cache.read_text().splitlines()) == 1
    # a line written before the set-up tile was tuned (three fields) applies the default tile
    cache.write_text(lines[0].split("\t")[0] + "\t" + " ".join(fields[:3]) + "\n")
    b2.set_setup_tile(6)
    assert b2.autotune(cache)[:2] == (shape, k) and b2.setup_tile() == 0


def test_autotune_falls_back_to_the_default_tile(gpu_available, tmp_path):
    """A cache path that is not a readable file gives the library default (32x32, K = 4) at once,
    without timing; a cache line that does not parse is skipped and the batch is tuned as usual.
    The flow is the same either way."""
    rows = [120, 120]
    say("autotune fallback on a 2-pair batch")
    b = Batch(rows)
    want = []
    for i, r in enumerate(rows):
        inputs, flow = single(r, 950 + i)
        b.set_inputs(i, *inputs)
        want.append(flow)
    b.set_tb_shape(0)
    b.set_iteration_kernel(6, 60 * 80)
    not_a_file = tmp_path / "cache_dir"
    not_a_file.mkdir()
    assert b.autotune(not_a_file) == (1, 4, 0.0)
    b.solve(graph=True)
    b.synchronize()
    for i in range(len(rows)):
        for a, w in zip(b.flow(i), want[i]):
            np.testing.assert_array_equal(a, w)
    # a corrupt line (unsupported shape and K) under this batch's key is not applied
    good = tmp_path / "probe.txt"
    b.autotune(good)
    key = good.read_text().splitlines()[0].split("\t")[0]
    bad = tmp_path / "bad.txt"
    bad.write_text(f"{key}\t7 13 0.5\n{key}\tgarbage\n")
    shape, k, ms = b.autotune(bad)
    assert shape in (0, 1) and k in (3, 4, 5, 6) and ms > 0
    assert bad.read_text().splitlines()[-1].split("\t")[1].split()[:2] == [str(shape), str(k)]


def test_full_size_batch_matches_single_pairs(gpu_available):
    """256 VGA pairs in one batch, the benchmark's size (a 141557760-thread iteration grid, about
    12 GB of planes): sampled pairs, the last one included, equal their single-pair solves."""
    from pdflow import synth_pairs_float
    n, rows = 256, 480
    say(f"{n} pairs of {rows}x640 in one batch")
    planes = synth_pairs_float(4200, n, rows, 640)
    b = Batch([rows] * n)
    for i in range(n):
        b.set_inputs(i, *(p[i] for p in planes))
    b.solve(graph=True)
    b.synchronize()
    g = Solver(rows, 1)
    for i in (0, 97, 200, n - 1):
        want = g.solve_pair(*(p[i] for p in planes))
        for a, w in zip(b.flow(i), want):
            np.testing.assert_array_equal(a, w, err_msg=f"pair {i}")


def test_async_inputs_and_flows_match_the_blocking_copies(gpu_available):
    """pdf_batch_upload_inputs_async / pdf_batch_download_flows_async from page-locked memory: new
    inputs queued after a solve take effect in the next solve, and the flows copied out equal the
    blocking per-pair copies of a batch given the same inputs by pdf_batch_set_inputs."""
    rows = [480, 240, 120, 60, 240]
    say(f"async copies, batch of {rows}")
    first = [single(r, 700 + i)[0] for i, r in enumerate(rows)]
    second = [single(r, 800 + i)[0] for i, r in enumerate(rows)]
    b, ref = Batch(rows), Batch(rows)
    for i in range(len(rows)):
        b.set_inputs(i, *first[i])
        ref.set_inputs(i, *second[i])
    b.solve(graph=True)
    ref.solve(graph=True)
    ref.synchronize()
    n_in = sum(4 * r * (r * 4 // 3) for r in rows)
    n_out = sum(3 * r * (r * 4 // 3) for r in rows)
    src, dst = PinnedBuffer(n_in), PinnedBuffer(n_out)
    src.array[:] = np.concatenate([np.ascontiguousarray(x, np.float32).ravel() for pair in second for x in pair])
    dst.array[:] = np.nan
    with pytest.raises(PDFlowError):
        b.upload_inputs_async(src.array[:-1])
    with pytest.raises(PDFlowError):
        b.download_flows_async(dst.array[1:])
    for rep in range(2):
        b.upload_inputs_async(src.array)
        b.solve(graph=True)
        b.download_flows_async(dst.array)
        b.synchronize()
        off = 0
        for i, r in enumerate(rows):
            P = r * (r * 4 // 3)
            for k, w in enumerate(ref.flow(i)):
                np.testing.assert_array_equal(dst.array[off + k * P: off + (k + 1) * P], w, err_msg=f"pair {i} plane {k} rep {rep}")
            off += 3 * P
        dst.array[:] = np.nan
    src.free()
    dst.free()


def test_every_setup_tile_matches_single_pairs(gpu_available):
    """The level-init and median-filter stencils of a batch give the same bits for every set-up
    tile (32x8 .. 8x8 pixels, 1 to 4 waves), on a batch whose levels are not multiples of the
    tiles (15x20, 30x40, ... 480x640)."""
    rows = [480, 240, 120, 60, 30, 15, 240]
    want, b = [], Batch(rows)
    for i, r in enumerate(rows):
        inputs, flow = single(r, 1300 + i)
        b.set_inputs(i, *inputs)
        want.append(flow)
    with pytest.raises(PDFlowError):
        b.set_setup_tile(len(Batch.SETUP_TILES))
    for tile, (tr, tc) in enumerate(Batch.SETUP_TILES):
        say(f"set-up tile {tr}x{tc}")
        b.set_setup_tile(tile)
        assert b.setup_tile() == tile
        b.solve(graph=tile % 2 == 0)
        b.synchronize()
        for i in range(len(rows)):
            for a, w in zip(b.flow(i), want[i]):
                np.testing.assert_array_equal(a, w, err_msg=f"tile {tr}x{tc}, pair {i}")


def test_async_copies_when_frames_are_larger_than_the_solve(gpu_available):
    """Frames larger than the solved level (cam_mode below 480 / rows): the asynchronous upload
    takes each pair's frames at their own size and the download gives the flow at the solved
    size, equal to the blocking copies."""
    from pdflow import synth_pair_float
    rows, cams = [120, 240, 120], [2, 1, 4]
    say(f"async copies, rows {rows}, cam_modes {cams}")
    frames = [synth_pair_float(1400 + i, 480 // c, 640 // c) for i, c in enumerate(cams)]
    b, ref = Batch(rows, cam_modes=cams), Batch(rows, cam_modes=cams)
    for i, f in enumerate(frames):
        ref.set_inputs(i, *f)
    ref.solve(graph=True)
    ref.synchronize()
    src = PinnedBuffer(sum(4 * (480 // c) * (640 // c) for c in cams))
    dst = PinnedBuffer(sum(3 * r * (r * 4 // 3) for r in rows))
    src.array[:] = np.concatenate([np.ascontiguousarray(x, np.float32).ravel() for f in frames for x in f])
    b.upload_inputs_async(src.array)
    b.solve(graph=True)
    b.download_flows_async(dst.array)
    b.synchronize()
    off = 0
    for i, r in enumerate(rows):
        P = r * (r * 4 // 3)
        for k, w in enumerate(ref.flow(i)):
            np.testing.assert_array_equal(dst.array[off + k * P: off + (k + 1) * P], w, err_msg=f"pair {i} plane {k}")
        off += 3 * P
    src.free()
    dst.free()


def test_two_batches_in_flight_ordered_by_wait(gpu_available):
    """pdf_batch_wait: two batches solved alternately, each solve queued behind the other batch's
    (the pipelined pattern of the benchmark, with a rule-like metrics pass and a re-solve on the
    current batch overlapping the next solve), give every pair's single-pair flow; waiting on
    itself is a no-op and a null batch is refused."""
    import ctypes
    rows = [240, 120, 60]
    say(f"two batches ordered by pdf_batch_wait, {rows}")
    pairs = [single(r, 1500 + i) for i, r in enumerate(rows)]
    a, b = Batch(rows), Batch(rows)
    for i, (frames, _) in enumerate(pairs):
        a.set_inputs(i, *frames)
        b.set_inputs(i, *frames)
    a.wait_for(a)
    cur, nxt = a, b
    cur.solve(graph=True)
    for step in range(4):
        nxt.wait_for(cur)
        nxt.solve(graph=True)
        cur.quality()                     # host waits on cur's stream only
        cur.resolve_plain([1])
        cur.synchronize()
        for i, (_, ref) in enumerate(pairs):
            for k in range(3):
                np.testing.assert_array_equal(cur.flow(i)[k], ref[k], err_msg=f"step {step} pair {i} plane {k}")
        cur, nxt = nxt, cur
    cur.synchronize()
    lib = a.lib
    assert lib.pdf_batch_wait(None, a.ctx) != 0
    assert lib.pdf_batch_wait(a.ctx, None) != 0
