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


@pytest.mark.parametrize("WS", [64, 128, 256, 512, 1024, 2048, 4096])
def test_batched_sync_staging_rounds(WS):
    """tile_block_sync_batched: ROUNDS = ceil(NF4 / BLOCK) loads per thread, only the last
    round partial; every chunk (tile g, row rr) loaded once, at source offset g*TR + rr."""
    for RB in (1, 2, 4):
        M = WS // 2
        G = (M + 1 + CW - 1) // CW
        NF4 = G * RB * CW // 2
        BLOCK = RB * (M // V)
        ROUNDS = (NF4 + BLOCK - 1) // BLOCK
        assert (ROUNDS - 1) * BLOCK < NF4 <= ROUNDS * BLOCK
        src = []
        for tid in range(BLOCK):
            for j in range(ROUNDS):
                f = tid + j * BLOCK
                if j + 1 < ROUNDS or f < NF4:
                    src.append((f // RB) * TR + f % RB)
        assert len(src) == NF4 == len(set(src))
        # all chunks lie inside the tile row's G tiles of TR chunks
        assert max(src) < G * TR
