"""batch_file (gt_locate_file: native decode into pinned buffers, chunked,
the next chunk read while the current one is located) must give the oracle's
rows on the frames that were written (by the independent encoders of
tests/movie_files.py), and exactly the rows of batch on the same frames:
TIFF (compressed and not), ETS (square and not, with execute_file's
row/column reading of the ETS frame size), and a movie of several chunks
linked across chunk boundaries."""
import numpy as np
import pytest

from movie_files import write_ets, write_tiff
from oracle import oracle
from tracking_optimizations_amd import _lib, synth
from tracking_optimizations_amd.gpu_tracking import batch_file_rust, batch_rust

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _lib.default_context(0)


@pytest.fixture(scope="module")
def frames512(ctx):
    return synth.generate_host(ctx, synth.single_512(seed=21, n_frames=4))


KW = dict(separation=9, minmass=500.0, percentile=64)


def _assert_oracle_rows(rows, frames, kw, diameter=9):
    """Rows (frame order, raster order of the candidate within a frame) equal
    the oracle's on the written frames: same count per frame, positions within
    1e-4 px, masses within 1e-5 relative."""
    p = oracle.Params(diameter=diameter, separation=kw["separation"], minmass=kw["minmass"],
                      percentile=kw["percentile"])
    ref = np.concatenate([oracle.locate_frame(f.astype(np.float32), p, frame_index=i)[0]
                          for i, f in enumerate(frames)])
    assert rows.shape == ref.shape, (rows.shape, ref.shape)
    assert np.array_equal(rows[:, 0], ref[:, 0])
    assert np.abs(rows[:, 2:4] - ref[:, 2:4]).max() <= 1e-4
    assert (np.abs(rows[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])).max() <= 1e-5


@pytest.mark.parametrize("comp,pred", [(1, 1), (5, 2), (8, 1)])
def test_batch_file_tiff(tmp_path, frames512, comp, pred):
    f = tmp_path / "movie.tif"
    write_tiff(f, list(frames512), rows_per_strip=64, compression=comp, predictor=pred)
    rows, cols = batch_file_rust(str(f), 9, **KW)
    _assert_oracle_rows(rows, frames512, KW)
    ref, rc = batch_rust(frames512, 9, **KW)
    assert cols == rc
    np.testing.assert_array_equal(rows, ref)


def test_batch_file_tiff_u8(tmp_path, frames512):
    v = np.clip(frames512 // 8, 0, 255).astype(np.uint8)
    f = tmp_path / "movie.tif"
    write_tiff(f, list(v))
    rows = batch_file_rust(str(f), 9, **KW)[0]
    _assert_oracle_rows(rows, v, KW)
    np.testing.assert_array_equal(rows, batch_rust(v, 9, **KW)[0])


@pytest.mark.parametrize("dims", [(512, 512), (384, 256)])
def test_batch_file_ets(tmp_path, ctx, dims):
    d0, d1 = dims
    spec = synth.single_512(seed=22, n_frames=3)
    v = synth.generate_host(ctx, spec)[:, :d1, :d0].copy()        # stored frames: d1 x d0
    f = tmp_path / "movie.ets"
    write_ets(f, list(v), (d0, d1), [(1, k, k) for k in range(3)])
    rows, _ = batch_file_rust(str(f), 9, channel=1, **KW)
    # execute_file reads the ETS frame size (dims[0], dims[1]) as (rows, columns)
    _assert_oracle_rows(rows, v.reshape(3, d0, d1), KW)
    ref, _ = batch_rust(v.reshape(3, d0, d1), 9, **KW)
    np.testing.assert_array_equal(rows, ref)


def test_batch_file_several_chunks_linked(tmp_path, ctx):
    spec = synth.movie_2048(seed=2, n_frames=40)       # 31 frames per 256 MiB chunk
    v = synth.generate_host(ctx, spec)
    f = tmp_path / "movie.tif"
    write_tiff(f, list(v))
    kw = dict(separation=11, smoothing_size=11, percentile=64, search_range=4.0, memory=2)
    rows, cols = batch_file_rust(str(f), 11, **kw)
    ref, rc = batch_rust(v, 11, **kw)
    assert cols == rc and rows.shape[0] > 40 * 1000
    np.testing.assert_array_equal(rows, ref)
