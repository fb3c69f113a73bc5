"""CPU checks of kernel index arithmetic (no GPU): every element a copy loop is meant
to move is moved exactly once and nothing outside the buffers is touched.

``tile_block_async`` (``leopard-em_amd/csrc/match.hip``) copies one row-pass tile block
of NF4 16-byte chunks into LDS: wave w starts at chunk w * WL, the block advances by
STEP = BLOCK chunks per round, lane l copies chunk base + l when it is below NF4.  The
loop bounds are compile-time constants of the launch (BLOCK = TR * (W/2) / 8 threads,
WL = min(BLOCK, 64) active lanes per wave).  An earlier form took the wave count from
BLOCK / 64, which is zero for sub-wave blocks (W = 128: 32 threads) and left the loop
without progress.  Sub-wave blocks are now also refused by
``tdtm_match_plan_set_option("rowpass_staging", 1 | 2)`` (W < 256).
"""

import pytest

TR, CW, V = 4, 2, 8


def _tile_chunks(WS):
    M = WS // 2
    G = (M + 1 + CW - 1) // CW
    TILE = G * TR * CW          # complex values staged per block
    return TILE // 2            # 16-byte chunks (two complex64 each)


def _async_copy_chunks(WS):
    """Chunks copied by each thread of tile_block_async<WS>, as the kernel indexes them."""
    NF4 = _tile_chunks(WS)
    BLOCK = TR * (WS // 2) // V
    WL = BLOCK if BLOCK < 64 else 64
    STEP = BLOCK
    assert BLOCK % WL == 0
    copied = []
    for tid in range(BLOCK):
        wave, lane = tid >> 6, tid & 63
        base = wave * WL
        rounds = 0
        while base < NF4:
            if base + lane < NF4:
                copied.append(base + lane)
            base += STEP
            rounds += 1
            assert rounds <= NF4, "copy loop does not advance"
    return NF4, copied


@pytest.mark.parametrize("WS", [64, 128, 256, 512, 1024, 2048, 4096])
def test_async_tile_copy_covers_every_chunk_once(WS):
    NF4, copied = _async_copy_chunks(WS)
    assert sorted(copied) == list(range(NF4))


def _tile_off(r, c):
    """tile_off: column-major 4-row x 2-column tiles of the intermediate."""
    return c * TR + r


@pytest.mark.parametrize("WS", [64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_batched_sync_staging_rounds(WS):
    """tile_block_sync_batched: ROUNDS = ceil(NE / BLOCK) 8-byte loads per thread, only the
    last round partial; every element (tile g, column c, row rr) loaded once, from
    g*TR*CW + tile_off(rr, c) of the tile row, and written once, to row rr's position 2g+c."""
    for RB in (1, 2, 4):
        M = WS // 2
        G = (M + 1 + CW - 1) // CW
        NE = G * RB * CW
        BLOCK = RB * (M // V)
        ROUNDS = (NE + BLOCK - 1) // BLOCK
        assert (ROUNDS - 1) * BLOCK < NE <= ROUNDS * BLOCK
        src, dst = [], []
        for tid in range(BLOCK):
            for j in range(ROUNDS):
                e = tid + j * BLOCK
                if j + 1 < ROUNDS or e < NE:
                    rr, c, g = e % RB, (e // RB) % CW, e // (RB * CW)
                    src.append(g * TR * CW + _tile_off(rr, c))
                    dst.append((rr, CW * g + c))
        assert len(src) == NE == len(set(src)) == len(set(dst))
        # within the tile row's G tiles, and only rows 0 .. RB-1 of each column
        assert max(src) < G * TR * CW and all(s % TR < RB for s in src)
        assert sorted(dst) == [(r, x) for r in range(RB) for x in range(CW * G)]


def _xs_slot(y, col, NCOL):
    """colpass_four's LDS slot (float2) of row y of the block's column col."""
    if NCOL == 1:
        return y
    return 8 * (y >> 2) + 4 * (col ^ (((y >> 3) ^ (y >> 4)) & 1)) + (y & 3)


def _xs_piece(u, NCOL):
    """The float4 slot thread u reads for piece u."""
    if NCOL == 1:
        return u
    trow = u // (TR // 2 * NCOL)
    return u ^ ((((trow >> 1) ^ (trow >> 2)) & 1) << 1)


@pytest.mark.parametrize("HH", [512, 1024, 2048, 4096])
@pytest.mark.parametrize("NCOL", [1, 2])
def test_column_pass_pieces_cover_tiles_once(HH, NCOL):
    """colpass_four's transposed stores: thread q (NCOL * T threads) writes pieces
    u = q + j NCOL T, j < HH / (2 T): tile row u / PPT, column (NCOL == 1: the block's
    column c), rows 2 h2, 2 h2 + 1 -- together every (row, column) of the block's columns
    once, 16-byte aligned, and 4 (2 with one column) consecutive threads per tile.  The LDS
    slots the rows are written to are distinct and within NCOL * HH float2, and the 16-byte
    slot a thread reads for its piece holds exactly that piece's two rows."""
    T = HH // V
    PPT = TR // 2 * NCOL
    cols_all = [0] if NCOL == 1 else list(range(CW))
    slot_of = {(y, cl): _xs_slot(y, cl, NCOL) for y in range(HH) for cl in cols_all}
    assert sorted(slot_of.values()) == list(range(NCOL * HH))
    for cl in cols_all:  # the writes' form: one slot per thread, immediate offsets
        for t in range(T):
            w = t if NCOL == 1 else 8 * (t >> 2) + 4 * (cl ^ (((t >> 3) ^ (t >> 4)) & 1)) + (t & 3)
            for i in range(V):
                assert w + (1 if NCOL == 1 else 2) * i * T == slot_of[(t + i * T, cl)]
    for c_block in range(CW if NCOL == 1 else 1):
        seen = []
        for q in range(NCOL * T):
            for j in range(HH // (2 * T)):
                u = q + j * NCOL * T
                trow = u // PPT
                col = c_block if NCOL == 1 else (u % PPT) // (TR // 2)
                h2 = u % (TR // 2)
                y = trow * TR + 2 * h2
                f4 = trow * 1000 * (TR * CW // 2) + (col * TR + 2 * h2) // 2  # tile row stride (any Gs)
                assert (col * TR + 2 * h2) % 2 == 0
                seen += [(y, col), (y + 1, col)]
                assert _tile_off(y % TR, col) // 2 == (col * TR + 2 * h2) // 2
                assert f4 >= 0
                lc = 0 if NCOL == 1 else col
                s4 = _xs_piece(u, NCOL)
                assert (slot_of[(y, lc)], slot_of[(y + 1, lc)]) == (2 * s4, 2 * s4 + 1)
                # the kernel's form: one slot per thread plus the loop's immediate offset
                rslot = q if NCOL == 1 else q ^ (((((q // PPT) >> 1) ^ ((q // PPT) >> 2)) & 1) << 1)
                assert rslot + j * NCOL * T == s4
        cols = [c_block] if NCOL == 1 else list(range(CW))
        assert sorted(seen) == sorted((y, c) for y in range(HH) for c in cols)


@pytest.mark.parametrize("HH", [512, 4096])
def test_column_pass_lds_writes_spread_over_banks(HH):
    """With two columns per block, a wave's 8-byte writes (64 consecutive rows of one
    column) take 32 distinct float2 banks per 32 lanes (64 four-byte banks) and 16 per 16
    lanes (32 banks)."""
    for col in range(CW):
        for y0 in range(0, HH, 64):
            sl = [_xs_slot(y0 + l, col, 2) for l in range(64)]
            for h in range(2):
                assert len({x % 32 for x in sl[32 * h:32 * h + 32]}) == 32
            for h in range(4):
                assert len({x % 16 for x in sl[16 * h:16 * h + 16]}) == 16


def test_mfma_pieces_cover_tiles_once():
    """k_colpass_mfma's stores (MF_BLOCK = 1024 threads, HH = 4096): every (row, column) of
    the tile column once."""
    HH, BLOCK = 4096, 1024
    seen = []
    for tid in range(BLOCK):
        for jj in range(HH // BLOCK):
            u = tid + jj * BLOCK
            trow, col, h2 = u // (TR // 2 * CW), (u // (TR // 2)) % CW, u % (TR // 2)
            y = trow * TR + 2 * h2
            seen += [(y, col), (y + 1, col)]
    assert sorted(seen) == sorted((y, c) for y in range(HH) for c in range(CW))

