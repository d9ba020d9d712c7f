"""The serving loop's device arena: first-fit placement, merging of freed ranges, the memory limit, and no driver
allocation once its blocks exist (host logic, with a stand-in for device buffers)."""
import pytest

from pharos_amd import serve


class FakeBlock:
    allocations = 0

    def __init__(self, nbytes, device=0):
        FakeBlock.allocations += 1
        self.nbytes = int(nbytes)
        self.base = 1 << 40
        self.freed = False

    def at(self, offset):
        return self.base + int(offset)

    def free(self):
        self.freed = True


@pytest.fixture
def arena(monkeypatch):
    monkeypatch.setattr(serve, "DeviceBuffer", FakeBlock)
    FakeBlock.allocations = 0
    return serve.DeviceArena(0, limit_bytes=1 << 20, block_bytes=256 << 10)


def test_first_fit_and_alignment(arena):
    a = arena.alloc(100)
    b = arena.alloc(5000)
    assert (a.offset, a.nbytes) == (0, 4096)
    assert (b.offset, b.nbytes) == (4096, 8192)
    assert b.ptr == a.ptr + 4096
    assert arena.used == 12288 and FakeBlock.allocations == 1


def test_freed_ranges_merge_and_are_reused(arena):
    bufs = [arena.alloc(4096) for _ in range(4)]
    bufs[1].free()
    bufs[2].free()
    assert arena.ranges[0][0] == [4096, 8192]          # the two neighbours became one range
    c = arena.alloc(8192)
    assert c.offset == 4096 and FakeBlock.allocations == 1
    for x in (bufs[0], bufs[3], c):
        x.free()
    assert arena.ranges[0] == [[0, 256 << 10]] and arena.used == 0


def test_limit_and_oversized_buffers(arena):
    big = arena.alloc(600 << 10)                        # larger than a block: a block of its own
    assert big is not None and arena.reserved() == (600 << 10) + 0
    small = arena.alloc(300 << 10)                      # a new 256 KiB block would exceed 1 MiB: an exact one
    assert small is not None and arena.reserved() == (900 << 10)
    assert arena.alloc(200 << 10) is None               # nothing left within the limit
    assert arena.alloc_all([4096, 200 << 10]) is None and arena.used == (900 << 10)  # all or nothing


def test_close_returns_blocks(arena):
    arena.alloc(4096)
    blocks = list(arena.blocks)
    arena.close()
    assert all(b.freed for b in blocks) and arena.used == 0 and not arena.blocks


def test_held_state_counts_against_the_limit(arena):
    # 1 MiB limit: holding 900 KiB for a batch's state leaves room for 124 KiB of buffers, no more.
    assert arena.hold(900 << 10)
    assert arena.alloc(200 << 10) is None
    b = arena.alloc(100 << 10)
    assert b is not None
    assert not arena.hold(100 << 10)
    arena.unhold(900 << 10)
    assert arena.hold(500 << 10) and arena.alloc(300 << 10) is not None


def test_request_state_share_covers_the_row_prepass():
    from pharos_amd.scenario import c3
    for sc in c3(seed=4, count=40):
        r = serve.Request(sc)
        rows = r.slice_epochs * len(sc.channels)
        prep = rows * serve.PREP_BYTES if sc.samples_per_epoch > 4096 else 0
        items = r.slice_epochs * -(-sc.samples_per_epoch // 4096) * serve.ITEM_BYTES
        # the share covers the pre-pass and the items' segments with the block's rounding (below 256 MiB a power
        # of two: at most 2x)
        assert r.state_bytes() >= 2 * (prep + items)
        assert r.state_bytes() <= 2 * (prep + items) + 64 * (r.slice_epochs // 100 + 2) * 2 + 1024
