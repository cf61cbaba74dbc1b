"""The launch rule of multi-step runs on periodic all-fluid lattices
(lbmx_launch_plan, host only): which kernel family a size and precision get,
the strip and segment shape of row streaming and the tile shape of the
three-step tiles."""
import pytest

from lbc_amd import native


def strip_cols(d, prec):
    """Output columns of a d-level strip: 64 - 2(d - 1), rounded down to whole
    32-byte sectors of a row (FP64 4 values, FP32 storage 8)."""
    q = 4 if prec == native.F64 else 8
    return (64 - 2 * (d - 1)) // q * q


@pytest.mark.parametrize("Nx,Ny,prec,depth", [
    (2048, 2048, native.F64, 5), (256, 256, native.F64, 3), (1024, 1024, native.F64, 4),
    (512, 512, native.F64, 3), (16384, 16384, native.F64, 5), (1536, 1536, native.F64, 4),
    (2048, 1800, native.F64, 5), (2048, 1600, native.F64, 4), (4096, 4096, native.F64, 5),
    (2048, 2048, native.F32_SHIFTED, 4), (1024, 1024, native.F32_SHIFTED, 4),
    (3072, 3072, native.F32_SHIFTED, 8), (768, 768, native.F32_SHIFTED, 4),
    (576, 576, native.F32_SHIFTED, 3), (640, 640, native.F32_SHIFTED, 4),
    (16384, 16384, native.F32_SHIFTED, 8), (2048, 2048, native.F32_WIDE, 4),
    (8192, 256, native.F64, 4), (64, 100000, native.F64, 5), (64, 30000, native.F64, 4),
    # around the streaming threshold (200 workgroups of 32 rows at four steps, 320
    # at eight, 400 for F32_WIDE)
    (576, 576, native.F64, 3), (640, 640, native.F64, 4), (704, 704, native.F64, 4),
    (768, 768, native.F64, 4), (768, 768, native.F32_WIDE, 3),
    (896, 896, native.F32_WIDE, 4), (1408, 1408, native.F32_SHIFTED, 4),
    (1536, 1536, native.F32_SHIFTED, 4), (896, 896, native.F32_SHIFTED, 4),
    (16384, 256, native.F32_SHIFTED, 4), (2448, 2448, native.F32_SHIFTED, 4),
    (2450, 2450, native.F32_SHIFTED, 8)])
def test_auto_rule(Nx, Ny, prec, depth):
    d, cols, rows = native.launch_plan(Nx, Ny, prec)
    assert d == depth
    if d == 5:
        # FP64 five steps: 56-column strips, 256 rows where that still gives 4,096
        # workgroups, else 128; only where 128-row segments give 512 workgroups
        assert prec == native.F64 and cols == 56
        strips = -(-Nx // cols)
        assert rows == min(Ny, 256 if strips * -(-Ny // 256) >= 4096 else 128)
        assert strips * -(-Ny // rows) >= 512
    elif d in (4, 8):
        # strip of 64 - 2(D - 1) columns rounded to 32-byte sectors; rows a power of
        # two in the depth's range
        assert cols == strip_cols(d, prec)
        strips = -(-Nx // cols)
        lo, hi = (32, 128) if d == 4 else (128, 512)
        if d == 4 and rows < 32:
            # 16-row segments only where 32-row ones give fewer than 768 workgroups,
            # lengthened to fit one round of 1,024 workgroups when 16 rows give up to
            # 1.5 times that many
            assert strips * -(-Ny // 32) < 768
            w16 = strips * -(-Ny // 16)
            assert rows == 16 or 1024 < w16 <= 1536
            base = 32
        else:
            assert rows == min(Ny, rows) and (rows & (rows - 1)) == 0 and lo <= rows <= hi
            assert not (d == 4 and rows == 32 and strips * -(-Ny // 32) < 768)
            base = rows
        # streamed only with >= 200 workgroups (at the base row count) at four
        # steps, 320 at eight, 400 with FP64 collision on FP32 storage; FP32 eight
        # steps from 6 M cells
        assert strips * -(-Ny // base) >= (400 if prec == native.F32_WIDE else 320 if d == 8
                                           else 200)
        if prec == native.F32_SHIFTED:
            assert (d == 8) == (Nx * Ny >= 6_000_000)
        if prec == native.F64:
            # four FP64 steps only where five would give fewer than 512 workgroups
            # (five levels: 56 columns)
            assert -(-Nx // 56) * -(-Ny // 128) < 512
    else:
        # tiles of 64 columns: 16 rows with FP64 collision on FP32 storage, else
        # the tallest of 8, 4, 2 rows that gives >= 1024 tiles (2 if none does)
        assert cols == 64
        if prec == native.F32_WIDE:
            assert rows == 16
        else:
            tiles = lambda ty: -(-Nx // 64) * -(-Ny // ty)
            assert rows in (8, 4, 2) and (tiles(rows) >= 1024 or rows == 2)
            assert rows == 8 or tiles(2 * rows) < 1024


@pytest.mark.parametrize("N,rows", [(256, 2), (300, 2), (512, 4), (600, 4), (768, 8), (1000, 8)])
def test_tile_rows(N, rows):
    for prec in (native.F64, native.F32_SHIFTED):
        assert native.launch_plan(N, N, prec, native.LAUNCH_TILES) == (3, 64, rows)
    assert native.launch_plan(N, N, native.F32_WIDE, native.LAUNCH_TILES) == (3, 64, 16)


@pytest.mark.parametrize("N,prec,rows", [
    (768, native.F64, 16), (1024, native.F64, 20), (1280, native.F64, 32),
    (1024, native.F32_SHIFTED, 20), (1280, native.F32_SHIFTED, 32),
    (1024, native.F32_WIDE, 20), (1280, native.F32_WIDE, 32)])
def test_small_stream_rows(N, prec, rows):
    assert native.launch_plan(N, N, prec) == (4, 56, rows)


def test_one_round_rows():
    # 16-row segments would give 1,152 workgroups (a second, partial round of the
    # 1,024 the CUs hold); longer ones fit one round
    for Nx, Ny in ((1024, 1024), (2048, 512), (1100, 1000)):
        d, cols, rows = native.launch_plan(Nx, Ny, native.F64)
        strips = -(-Nx // cols)
        assert strips * -(-Ny // 16) > 1024
        assert strips * -(-Ny // rows) <= 1024 < strips * -(-Ny // (rows - 1))


def test_headline_shape():
    # 16384^2 FP64: 293 strips of 56 columns x 64 segments of 256 rows
    assert native.launch_plan(16384, 16384, native.F64) == (5, 56, 256)
    assert native.launch_plan(4096, 4096, native.F64) == (5, 56, 128)
    # FP32: eight levels, 50 columns rounded to 48 (192-byte rows, six sectors)
    assert native.launch_plan(16384, 16384, native.F32_SHIFTED) == (8, 48, 512)
    # forced streaming keeps eight FP32 steps at any size
    assert native.launch_plan(1024, 1024, native.F32_SHIFTED, native.LAUNCH_STREAM)[0] == 8


def test_forced_modes():
    for prec, sd in ((native.F64, 5), (native.F32_SHIFTED, 8), (native.F32_WIDE, 4)):
        assert native.launch_plan(300, 200, prec, native.LAUNCH_SINGLE) == (1, 0, 0)
        assert native.launch_plan(16384, 16384, prec, native.LAUNCH_TILES) == \
            (3, 64, 16 if prec == native.F32_WIDE else 8)
        d, cols, rows = native.launch_plan(3, 150, prec, native.LAUNCH_STREAM)
        assert d == sd and cols == strip_cols(sd, prec) and rows <= 150


def test_bad_arguments():
    with pytest.raises(native.LbmxError):
        native.launch_plan(2, 100)
    with pytest.raises(native.LbmxError):
        native.launch_plan(100, 100, 7)
    with pytest.raises(native.LbmxError):
        native.launch_plan(100, 100, native.F64, 9)
