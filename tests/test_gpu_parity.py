"""GPU locate vs the CPU oracle on the same frames (bit-exact candidate sets,
positions within 1e-4 px, mass and characterisation within 1e-5 relative).

Frames come from the device generator (copied back for the oracle), from the
reference's own grey_lion frame (tests/golden) and from crops of odd sizes
that exercise partial tiles and frame borders.
"""
import numpy as np
import pytest

from oracle import oracle
from tracking_optimizations_amd import _lib, synth
from tracking_optimizations_amd.gpu_tracking import locate, make_params

pytestmark = pytest.mark.gpu

POS_TOL = 1e-4     # px, on y and x
REL_TOL = 1e-5     # mass and characterisation columns


@pytest.fixture(scope="module")
def ctx():
    return _lib.default_context(0)


@pytest.fixture(scope="module")
def frames512(ctx):
    return synth.generate_host(ctx, synth.single_512(seed=1, n_frames=3))


@pytest.fixture(scope="module")
def grey_lion():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "grey_lion_u8.npy"))


CASES = {
    "c1": dict(diameter=9, minmass=1000, separation=9, percentile=64),
    "c2": dict(diameter=11, minmass=0, separation=11, smoothing_size=11, percentile=64),
    "defaults": dict(diameter=9),
    "char": dict(diameter=9, separation=9, characterize=True, cpu_processed=True),
    "even_d": dict(diameter=10, separation=12, smoothing_size=7),
    "iter1": dict(diameter=9, separation=9, max_iterations=1),
    "iter0": dict(diameter=9, separation=9, max_iterations=0),
    "small_sep": dict(diameter=7, separation=5, smoothing_size=9),
    "wide_smooth": dict(diameter=9, separation=9, smoothing_size=17, noise_size=2.0),
    "wide_sep": dict(diameter=9, separation=20),
    # the default separation of diameter 13 (14): dilation window 19, the
    # row-first detect kernel (windows wider than the column-first one's lanes)
    "d13_default_sep": dict(diameter=13),
    # window 22 (diameter 15, default separation 16), and windows 25 and 26
    # with smoothing radius 8 and disk radius 8 (diameter 17), all with the
    # walk fused into detect
    "d15_default_sep": dict(diameter=15),
    "d17_default_sep": dict(diameter=17),
    "d17_sep19": dict(diameter=17, separation=19, percentile=64),
    "pct0": dict(diameter=9, separation=9, percentile=0),
    "pct100": dict(diameter=9, separation=9, percentile=100),
    # dilation halo (3) narrower than the disk radius (5): walks near tile
    # edges read their disks from P instead of the staged tile
    "halo_fallback": dict(diameter=11, separation=5, smoothing_size=9, percentile=50),
    # dilation window of 70 px: too wide for a staged tile, the row maxima go
    # through a plane
    "sep_50": dict(diameter=9, separation=50, percentile=64),
}


def _oracle_params(kw):
    return oracle.Params(
        diameter=kw["diameter"], minmass=kw.get("minmass", 0.0),
        separation=kw.get("separation"), noise_size=kw.get("noise_size", 1.0),
        smoothing_size=kw.get("smoothing_size"), max_iterations=kw.get("max_iterations", 10),
        percentile=kw.get("percentile"), characterize=kw.get("characterize", False),
        cpu_processed=kw.get("cpu_processed", False))


def _oracle_movie(frames, kw):
    rows, cands = [], []
    for i, f in enumerate(frames):
        r, c, _ = oracle.locate_frame(f.astype(np.float32), _oracle_params(kw), frame_index=i)
        rows.append(r)
        cands.append(c)
    ncol = oracle.ncols(_oracle_params(kw))
    return (np.concatenate(rows) if rows else np.zeros((0, ncol)),
            np.concatenate(cands) if cands else np.zeros(0, np.int64))


def _check(rows, cand, ref_rows, ref_cand, label):
    assert rows.shape[1] == ref_rows.shape[1], label
    o = np.lexsort((cand, rows[:, 0]))
    r = np.lexsort((ref_cand, ref_rows[:, 0]))
    rows, cand = rows[o].astype(np.float64), cand[o]
    ref_rows, ref_cand = ref_rows[r], ref_cand[r]
    assert len(cand) == len(ref_cand), f"{label}: {len(cand)} rows vs oracle {len(ref_cand)}"
    assert np.array_equal(cand, ref_cand), f"{label}: candidate pixel sets differ"
    assert np.array_equal(rows[:, 0], ref_rows[:, 0]), f"{label}: frame column differs"
    if len(cand) == 0:
        return
    dpos = np.abs(rows[:, 2:4] - ref_rows[:, 2:4]).max()
    assert dpos <= POS_TOL, f"{label}: position error {dpos}"
    rel = np.abs(rows[:, 1] - ref_rows[:, 1]) / np.maximum(np.abs(ref_rows[:, 1]), 1e-30)
    assert rel.max() <= REL_TOL, f"{label}: mass rel error {rel.max()}"
    if rows.shape[1] > 4:
        a, b = rows[:, 4:], ref_rows[:, 4:]
        # the same non-finite entries on both sides, the finite ones close
        assert np.array_equal(np.isnan(a), np.isnan(b)), f"{label}: NaN pattern differs"
        assert np.array_equal(np.isinf(a), np.isinf(b)), f"{label}: inf pattern differs"
        fin = np.isfinite(b)
        err = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1.0)
        assert err.size == 0 or err.max() <= REL_TOL, f"{label}: characterisation rel error {err.max()}"


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("plain", [False, True])
def test_synthetic_u16(ctx, frames512, case, plain):
    kw = CASES[case]
    print(f"[parity] synthetic u16 case={case} plain={plain}", flush=True)
    rows, cand, cols, _ = locate(frames512, make_params(**kw), plain=plain, ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(frames512, kw)
    _check(rows, cand, ref_rows, ref_cand, f"u16/{case}/plain={plain}")


@pytest.mark.parametrize("case", ["c1", "defaults", "char", "even_d", "wide_smooth"])
def test_float32_input(ctx, frames512, case):
    kw = CASES[case]
    f32 = frames512[:1].astype(np.float32) * np.float32(0.75)
    print(f"[parity] float32 case={case}", flush=True)
    rows, cand, _, _ = locate(f32, make_params(**kw), ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(f32, kw)
    _check(rows, cand, ref_rows, ref_cand, f"f32/{case}")


def _float_frames(frames, seed):
    """float32 frames whose values carry full 24-bit mantissas and some
    negative values: boxcar sums of them round differently when taken in
    another order (checked here on the CPU)."""
    rng = np.random.default_rng(seed)
    f = (frames.astype(np.float32) * np.float32(0.7310585)
         + rng.random(frames.shape, dtype=np.float32) * np.float32(3.0) - np.float32(1.0))
    f[:, ::37, :] *= np.float32(-0.5)   # rows of negative values
    # a running boxcar (add the entering tap, drop the leaving one) differs
    # from the tap-by-tap sum on some pixels of these frames
    row = f[0, 7]
    k = 11
    direct = [np.float32(0.0)] * (len(row) - k + 1)
    for j in range(len(direct)):
        acc = np.float32(0.0)
        for i in range(k):
            acc = np.float32(acc + row[j + i])
        direct[j] = acc
    run = [direct[0]]
    for j in range(1, len(direct)):
        run.append(np.float32(np.float32(run[-1] + row[j + k - 1]) - row[j - 1]))
    assert any(a != b for a, b in zip(direct, run))
    return np.ascontiguousarray(f)


@pytest.mark.parametrize("case", ["c1", "defaults", "wide_smooth", "d13_default_sep", "iter1"])
def test_float32_rounding(ctx, frames512, case):
    """float32 input with rounding in its boxcar sums, on the streaming
    bandpass (widths that are multiples of 4: 512, and 100 with strips at
    both frame edges) and the generic one (width 98)."""
    kw = CASES[case]
    f = _float_frames(frames512[:2], 5)
    for sub in (f, f[:, :100, :100], f[:, :90, :98]):
        sub = np.ascontiguousarray(sub)
        print(f"[parity] float32 rounding case={case} shape={sub.shape}", flush=True)
        rows, cand, _, _ = locate(sub, make_params(**kw), ctx=ctx)
        ref_rows, ref_cand = _oracle_movie(sub, kw)
        _check(rows, cand, ref_rows, ref_cand, f"f32r/{case}/{sub.shape}")


@pytest.mark.parametrize("scale", [1e-33, 1e-25, 1e30, 1e33])
def test_float32_extreme_magnitudes(ctx, frames512, scale):
    """float32 frames of very small and very large magnitudes (sums from
    about 1e-31 to 1e38): the float32 bandpass keeps the shader's exact
    divisions and sums at every scale."""
    kw = dict(diameter=9, separation=9, minmass=0.0, percentile=64)
    f = np.ascontiguousarray((_float_frames(frames512[:1], 7) * np.float32(scale)).astype(np.float32))
    f = np.ascontiguousarray(f[:, :200, :256])
    print(f"[parity] float32 extreme magnitudes scale={scale}", flush=True)
    rows, cand, _, _ = locate(f, make_params(**kw), ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(f, kw)
    _check(rows, cand, ref_rows, ref_cand, f"f32x/{scale}")


@pytest.mark.parametrize("case", ["c1", "defaults", "char", "small_sep"])
def test_golden_grey_lion_u8(ctx, grey_lion, case):
    kw = CASES[case]
    print(f"[parity] grey_lion u8 case={case}", flush=True)
    rows, cand, _, _ = locate(grey_lion[None], make_params(**kw), ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(grey_lion[None], kw)
    _check(rows, cand, ref_rows, ref_cand, f"grey_lion/{case}")


@pytest.mark.parametrize("shape", [(100, 130), (65, 257), (300, 64), (17, 23), (5, 5)])
def test_odd_shapes(ctx, frames512, shape):
    h, w = shape
    crop = np.ascontiguousarray(frames512[:2, 37:37 + h, 11:11 + w])
    for case in ("c1", "defaults", "char"):
        kw = CASES[case]
        print(f"[parity] odd shape {shape} case={case}", flush=True)
        rows, cand, _, _ = locate(crop, make_params(**kw), ctx=ctx)
        ref_rows, ref_cand = _oracle_movie(crop, kw)
        _check(rows, cand, ref_rows, ref_cand, f"shape{shape}/{case}")


def test_deterministic_and_plain_equals_batched(ctx, frames512):
    p = make_params(**CASES["c2"])
    a = locate(frames512, p, ctx=ctx)
    b = locate(frames512, p, ctx=ctx)
    c = locate(frames512, p, plain=True, ctx=ctx)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_empty_movie(ctx):
    rows, cand, cols, _ = locate(np.zeros((0, 64, 64), np.uint16), make_params(9), ctx=ctx)
    assert rows.shape == (0, 4) and len(cand) == 0
    assert [c[0] for c in cols] == ["frame", "mass", "y", "x"]


def test_blank_frames(ctx):
    rows, cand, _, _ = locate(np.full((2, 128, 128), 7, np.uint16), make_params(9), ctx=ctx)
    assert len(rows) == 0


@pytest.mark.parametrize("per_launch", [1, 2])
def test_multi_batch_two_streams_identical(frames512, per_launch):
    """Small launches force several batches on alternating streams; rows (and
    their order) must not change."""
    p = make_params(**CASES["c2"])
    ref = locate(frames512, p, ctx=_lib.Context(0))
    ctx = _lib.Context(0)
    ctx.set_batch(per_launch)
    got = locate(frames512, p, ctx=ctx)
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("case", ["c1", "c2", "defaults", "even_d", "iter1", "small_sep",
                                  "wide_sep", "pct0", "pct100", "halo_fallback"])
def test_walk_in_detect_equals_refine_kernel(frames512, case):
    """The centroid walk fused into detect and the separate refine kernel
    give identical rows."""
    p = make_params(**CASES[case])
    fused = _lib.Context(0)
    separate = _lib.Context(0)
    separate.set_fused_walk(False)
    print(f"[parity] fused walk vs refine kernel case={case}", flush=True)
    a = locate(frames512, p, ctx=fused)
    b = locate(frames512, p, ctx=separate)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_graph_replay_matches_direct_launches(frames512):
    """Replays of the captured launch sequence, with direct launches and
    synchronisations in between, give the same rows every time."""
    p = make_params(**CASES["c2"])
    movie = np.ascontiguousarray(np.concatenate([frames512] * 3))
    ref = locate(movie, p, ctx=_lib.Context(0))
    gctx = _lib.Context(0)
    gctx.set_batch(2)           # several batches on both streams
    gctx.set_graphs(True)
    other = _lib.Context(0)
    for rep in range(4):
        print(f"[parity] graph replay {rep}", flush=True)
        got = locate(movie, p, ctx=gctx)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), rep
        locate(frames512[:1], make_params(**CASES["c1"]), ctx=other)  # eager work between
        other.synchronize()


def test_autotune_keeps_results(frames512):
    p = make_params(**CASES["c2"])
    ref = locate(frames512, p, ctx=_lib.Context(0))
    ctx = _lib.Context(0)
    t = ctx.autotune(p, 512, 512, n_frames=8)
    assert t["seg_rows"] in (64, 128, 256) and t["frames_per_launch"] in (0, 16, 32)
    assert t["fused_walk"] in (True, False)
    got = locate(frames512, p, ctx=ctx)
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])


def test_uneven_batch_matches_oracle(ctx, frames512):
    """gt_locate_batch with frames of different shapes in one call: every
    frame's rows equal the oracle's, and frame columns carry the frame index
    given in the descriptor."""
    import ctypes
    from tracking_optimizations_amd._lib import GtFrameDesc, Result, check, load
    shapes = [(256, 256), (100, 132), (512, 512), (64, 200), (300, 64), (256, 256)]
    frames = [np.ascontiguousarray(frames512[i % 3, :h, :w]) for i, (h, w) in enumerate(shapes)]
    kw = CASES["c1"]
    descs = (GtFrameDesc * len(frames))()
    for i, f in enumerate(frames):
        descs[i].data = f.ctypes.data
        descs[i].h, descs[i].w = f.shape
        descs[i].frame_index = 10 + i
    out = ctypes.c_void_p()
    p = make_params(**kw)
    print("[parity] uneven batch", flush=True)
    check(load().gt_locate_batch(ctx.handle, descs, len(frames), 0, ctypes.byref(p), 0,
                                 ctypes.byref(out)))
    r = Result(out)
    rows, cand = r.rows(), r.candidates()
    for i, f in enumerate(frames):
        sel = rows[:, 0] == 10 + i
        ref_rows, ref_cand, _ = oracle.locate_frame(f.astype(np.float32), _oracle_params(kw),
                                                    frame_index=10 + i)
        _check(rows[sel], cand[sel], ref_rows, ref_cand, f"uneven/{i}/{f.shape}")


@pytest.mark.parametrize("case", ["c1", "c2", "char", "small_sep"])
@pytest.mark.parametrize("plain", [False, True])
def test_row_order_matches_reference(ctx, frames512, case, plain):
    """Rows come in frame order and, within a frame, in raster order of the
    candidate pixel, as get_work (execute_gpu.rs:108-195) emits them: the
    unsorted output equals the oracle's row for row."""
    kw = CASES[case]
    rows, cand, _, _ = locate(frames512, make_params(**kw), plain=plain, ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(frames512, kw)
    assert np.array_equal(cand, ref_cand)
    assert np.array_equal(rows[:, 0], ref_rows[:, 0])
    assert np.abs(rows[:, 2:4] - ref_rows[:, 2:4]).max() <= POS_TOL


def test_full_size_c2_frames(ctx):
    """Two frames of the benchmark movie itself (2048x2048 uint16, 5000 spots,
    d = 11, separation 11, smoothing 11, percentile 64), device-resident as in
    the benchmark: bit-exact candidates and row order, positions and masses
    within the tolerances, on the batched and the plain path."""
    import ctypes
    spec = synth.movie_2048(seed=2, n_frames=2)
    movie = synth.generate(ctx, spec)
    host = movie.to_host(np.uint16, (2, 2048, 2048))
    kw = CASES["c2"]
    p = make_params(**kw)
    ref_rows, ref_cand = _oracle_movie(host, kw)
    print(f"[parity] full-size c2: oracle {len(ref_cand)} rows", flush=True)
    for flags in (_lib.GT_FRAMES_ON_DEVICE, _lib.GT_FRAMES_ON_DEVICE | _lib.GT_PLAIN):
        out = ctypes.c_void_p()
        _lib.check(_lib.load().gt_locate_u16(ctx.handle, ctypes.c_void_p(movie.ptr), 2, 2048,
                                             2048, ctypes.byref(p), flags, ctypes.byref(out)))
        r = _lib.Result(out)
        rows, cand = r.rows(), r.candidates()
        assert np.array_equal(cand, ref_cand), "row order or candidate set differs"
        _check(rows, cand, ref_rows, ref_cand, f"full-size c2 flags {flags}")
    movie.free()
    assert len(ref_cand) > 20000


def test_uneven_batch_full_shapes(ctx):
    """The shapes of the uneven benchmark batch (2048x2048, 512x1024, 256x256)
    in one gt_locate_batch call from host memory (staged through the pinned
    chunks): every frame equals the oracle."""
    import ctypes
    from tracking_optimizations_amd._lib import GtFrameDesc, Result, check, load
    big = synth.generate_host(ctx, synth.movie_2048(seed=3, n_frames=2))
    frames = [big[0], np.ascontiguousarray(big[1, 100:612, 300:1324]),
              np.ascontiguousarray(big[0, 700:956, 900:1156]), big[1],
              np.ascontiguousarray(big[1, 1500:1756, 17:273])]
    kw = CASES["c2"]
    descs = (GtFrameDesc * len(frames))()
    for i, f in enumerate(frames):
        descs[i].data = f.ctypes.data
        descs[i].h, descs[i].w = f.shape
        descs[i].frame_index = i
    out = ctypes.c_void_p()
    p = make_params(**kw)
    print("[parity] uneven full shapes", flush=True)
    check(load().gt_locate_batch(ctx.handle, descs, len(frames), 0, ctypes.byref(p), 0,
                                 ctypes.byref(out)))
    r = Result(out)
    rows, cand = r.rows(), r.candidates()
    for i, f in enumerate(frames):
        sel = rows[:, 0] == i
        ref_rows, ref_cand, _ = oracle.locate_frame(f.astype(np.float32), _oracle_params(kw),
                                                    frame_index=i)
        _check(rows[sel], cand[sel], ref_rows, ref_cand, f"uneven-full/{i}/{f.shape}")


def test_host_result_blocks_reused(ctx):
    """Host results of 1 MiB and more are kept after they are freed and reused
    by the next result that fits: a result written into a reused block (here
    four 2048x2048 frames with the characterisation columns, about 1.6 MB of
    rows, after a slightly larger result was freed) equals the oracle, and a
    repeat of the same call into the block freed by the first is identical."""
    spec = synth.movie_2048(seed=2, n_frames=4)
    host = synth.generate_host(ctx, spec)
    big = dict(diameter=11, separation=11, smoothing_size=11, percentile=64, characterize=True)
    kw = dict(diameter=11, separation=13, smoothing_size=11, percentile=64, characterize=True)
    r0 = locate(host, make_params(**big), ctx=ctx)[0]
    assert r0.nbytes >= (1 << 20)
    del r0
    rows, cand, _, _ = locate(host, make_params(**kw), ctx=ctx)
    assert rows.nbytes >= (1 << 20)
    ref_rows, ref_cand = _oracle_movie(host, kw)
    _check(rows, cand, ref_rows, ref_cand, "reused host blocks")
    rows2, cand2, _, _ = locate(host, make_params(**kw), ctx=ctx)
    assert np.array_equal(rows, rows2) and np.array_equal(cand, cand2)


TILE_H, TILE_W = 64, 48   # detect tile (rows, columns); dedup works on 2 x 2 blocks of them


def _max_staged_records(frame, kw):
    """Largest number of pre-dedup records any dedup block (2 x 2 tiles of
    64 rows x 48 columns) has to stage: the valid records of the tiles within
    its window (separation + 2 (max_iter + rint) + 1 px, rounded out to
    whole tiles past the block) that lie within separation + 1 of the
    bounding box of the block's own records; and the largest number of
    window tiles.  Computed on the CPU from the oracle's candidates."""
    p = _oracle_params(kw)
    cand, y, x, m = oracle.candidates_frame(frame.astype(np.float32), p)
    H, W = frame.shape
    g = oracle.geometry(p)
    sep = p.c().separation
    D = np.ceil(sep) + 2.0 * (p.max_iterations + g.rint) + 1.0
    rty, rtx = int(D // TILE_H) + 1, int(D // TILE_W) + 1
    nty, ntx = (H + TILE_H - 1) // TILE_H, (W + TILE_W - 1) // TILE_W
    ty, tx = (cand // W) // TILE_H, (cand % W) // TILE_W
    by, bx = ty // 2, tx // 2
    best, ntl_max = 0, 0
    for a in np.unique(by):
        for b in np.unique(bx[by == a]):
            own = (by == a) & (bx == b)
            y_lo, y_hi = max(0, 2 * a - rty), min(nty - 1, 2 * a + 1 + rty)
            x_lo, x_hi = max(0, 2 * b - rtx), min(ntx - 1, 2 * b + 1 + rtx)
            win = (ty >= y_lo) & (ty <= y_hi) & (tx >= x_lo) & (tx <= x_hi)
            ntl_max = max(ntl_max, (y_hi - y_lo + 1) * (x_hi - x_lo + 1))
            mg = sep + 1.0
            box = ((y >= y[own].min() - mg) & (y <= y[own].max() + mg) &
                   (x >= x[own].min() - mg) & (x <= x[own].max() + mg))
            best = max(best, int((win & box).sum()))
    return best, ntl_max


def test_dense_dedup_stages_in_rounds(ctx):
    """A noisy frame with a small disk and separation gives hundreds of
    records per dedup neighbourhood, more than one staging round holds (256):
    rows still equal the oracle."""
    rng = np.random.default_rng(21)
    frame = rng.poisson(300, (1, 384, 320)).astype(np.uint16)
    for kw in (dict(diameter=3, separation=3, smoothing_size=3),      # ~970 records: 4 rounds
               dict(diameter=3, separation=2, smoothing_size=5)):     # ~4600: 19 rounds
        staged, _ = _max_staged_records(frame[0], kw)
        print(f"[parity] dense dedup {kw}: up to {staged} staged records per tile", flush=True)
        assert staged > 256, staged
        ref_rows, ref_cand = _oracle_movie(frame, kw)
        for plain in (False, True):
            rows, cand, _, _ = locate(frame, make_params(**kw), plain=plain, ctx=ctx)
            _check(rows, cand, ref_rows, ref_cand, f"dense dedup {kw} plain={plain}")


@pytest.mark.parametrize("kw", [dict(diameter=9, separation=9, max_iterations=100),
                                dict(diameter=9, separation=200, percentile=64)],
                         ids=["max_iterations_100", "separation_200"])
def test_wide_dedup_window(ctx, kw):
    """Dedup windows of more than 64 tiles (separation + 2 (max_iterations +
    rint) + 1 of about 200 px and more) on 1024x1024 frames: the neighbour
    tiles are taken in groups of 64, and rows equal the oracle."""
    spec = synth.link_1024(seed=5, n_frames=2)
    frames = synth.generate_host(ctx, spec)
    _, ntl = _max_staged_records(frames[0], kw)
    print(f"[parity] wide dedup window {kw}: {ntl} neighbour tiles", flush=True)
    assert ntl > 64, ntl
    rows, cand, _, _ = locate(frames, make_params(**kw), ctx=ctx)
    ref_rows, ref_cand = _oracle_movie(frames, kw)
    _check(rows, cand, ref_rows, ref_cand, f"wide window {kw}")


def _boundary_frames(w, h=130, seed=48):
    """One uint16 frame of width w with spots centred on and beside the
    48-column tile boundaries and the 64-row one, plus two at the right
    edge, on a Poisson background."""
    rng = np.random.default_rng(seed + w)
    yy, xx = np.mgrid[0:h, 0:w]
    img = rng.poisson(80, (h, w)).astype(np.float64)
    for cx in (20.0, 47.0, 48.0, 71.5, 95.5, 96.0, w - 1.0, w - 3.5):
        for cy in (10.0, 63.5, 64.0, 100.0):
            img += 900.0 * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 1.6 ** 2))
    return np.floor(img).astype(np.uint16)[None]


@pytest.mark.parametrize("w", [97, 113, 143])
def test_tile_column_boundaries(ctx, w):
    """Detect tiles are 48 columns wide, their staged columns (tile and
    window halo) the 64 lanes of a wave: spots on and beside the tile
    boundaries, frames leaving 1, 17 and 47 columns in the last tile, and
    window sizes 7, 15 (column-first kernel) and 19 (row-first kernel) give
    the oracle's candidates and rows."""
    frames = _boundary_frames(w)
    for kw in (dict(diameter=7, separation=5),
               dict(diameter=11, separation=11, smoothing_size=11, percentile=64),
               dict(diameter=13, separation=14, percentile=64)):
        print(f"[parity] tile boundaries w={w} {kw}", flush=True)
        ref_rows, ref_cand = _oracle_movie(frames, kw)
        assert len(ref_cand) > 0
        for plain in (False, True):
            rows, cand, _, _ = locate(frames, make_params(**kw), plain=plain, ctx=ctx)
            _check(rows, cand, ref_rows, ref_cand, f"tile boundaries w={w} {kw} plain={plain}")


def test_uneven_workload_content(ctx):
    """Frames of the uneven benchmark batch itself (its generator: one spot per
    800 px^2, sigma U(1.2, 2.5), lognormal peaks, background U(50, 300) per
    frame; 256x256, 512x1024 and 2048x2048 shuffled), device-resident in one
    gt_locate_batch call, batched and plain: every frame equals the oracle."""
    import ctypes
    from tracking_optimizations_amd._lib import Result, check, load
    nf = 20
    movie = synth.UnevenMovie(ctx, nf, seed=3)
    descs = movie.frame_descs()
    kw = CASES["c2"]
    p = make_params(**kw)
    host = []
    for (ptr, h, w) in movie.descs:
        f = np.empty((h, w), np.uint16)
        check(load().gt_memcpy_d2h(ctx.handle, f.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.c_void_p(ptr), f.nbytes))
        host.append(f)
    assert {f.shape for f in host} == {(256, 256), (512, 1024), (2048, 2048)}
    refs = [oracle.locate_frame(f.astype(np.float32), _oracle_params(kw), frame_index=i)
            for i, f in enumerate(host)]
    for flags in (_lib.GT_FRAMES_ON_DEVICE, _lib.GT_FRAMES_ON_DEVICE | _lib.GT_PLAIN):
        out = ctypes.c_void_p()
        print(f"[parity] uneven workload content flags {flags}", flush=True)
        check(load().gt_locate_batch(ctx.handle, descs, nf, 0, ctypes.byref(p), flags,
                                     ctypes.byref(out)))
        r = Result(out)
        rows, cand = r.rows(), r.candidates()
        for i, (ref_rows, ref_cand, _) in enumerate(refs):
            sel = rows[:, 0] == i
            _check(rows[sel], cand[sel], ref_rows, ref_cand, f"uneven content {i} {host[i].shape}")
    movie.free()


def test_graph_replay_host_chunks(ctx):
    """Host frames of 40 x 2048^2 pass through the two pinned chunks (31
    frames each), so consecutive chunks differ in their staging buffer: with
    graphs on, each chunk's captured sequence is cached and replayed on the
    next call, and every replay gives the rows of direct launches."""
    host = synth.generate_host(ctx, synth.movie_2048(seed=2, n_frames=40))
    p = make_params(**CASES["c2"])
    ref = locate(host, p, ctx=_lib.Context(0))
    g = _lib.Context(0)
    g.set_graphs(True)
    for rep in range(3):
        print(f"[parity] graph replay of host chunks {rep}", flush=True)
        got = locate(host, p, ctx=g)
        assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), rep


@pytest.mark.parametrize("shift", [(0, 1), (3, -2), (17, 29)])
def test_full_size_translation(ctx, shift):
    """Two 1984x1984 windows of a 2048x2048 benchmark-movie frame, offset by
    `shift`: rows whose candidate lies more than 40 px inside both windows
    are the same rows, moved by the offset (candidate pixel exactly, mass bit
    for bit, position to binary32 rounding) - a size-independent check of the
    whole path at full size, without the oracle."""
    dy, dx = shift
    frame = synth.generate_host(ctx, synth.movie_2048(seed=2, n_frames=1))[0]
    n = 1984
    a = np.ascontiguousarray(frame[24:24 + n, 24:24 + n])
    b = np.ascontiguousarray(frame[24 + dy:24 + dy + n, 24 + dx:24 + dx + n])
    assert a.shape == b.shape == (n, n)
    # the c2 filters and separation without its percentile: the threshold is a
    # quantile of the whole frame and would differ between the windows
    p = make_params(diameter=11, minmass=0, separation=11, smoothing_size=11)
    ra, ca, _, _ = locate(a[None], p, ctx=ctx)
    rb, cb, _, _ = locate(b[None], p, ctx=ctx)
    m = 40
    inside = lambda c, oy, ox: ((c // n - oy >= m) & (c // n - oy < n - m) &
                                (c % n - ox >= m) & (c % n - ox < n - m))
    ka = inside(ca, 0, 0) & inside(ca, dy, dx)
    kb = inside(cb, -dy, -dx) & inside(cb, 0, 0)
    print(f"[parity] translation {shift}: {ka.sum()} rows compared", flush=True)
    assert ka.sum() > 10000
    assert np.array_equal(ca[ka] - (dy * n + dx), cb[kb])
    assert np.array_equal(ra[ka, 1], rb[kb, 1])
    # positions below 2048 carry a binary32 ulp of 1.2e-4 in each window
    assert np.abs(ra[ka, 2] - dy - rb[kb, 2]).max() <= 2.5e-4
    assert np.abs(ra[ka, 3] - dx - rb[kb, 3]).max() <= 2.5e-4
