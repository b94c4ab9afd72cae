"""The serving loop's job-state allocator, run on the host (no GPU) against a
plain model of first-fit allocation with coalescing of free neighbours."""

import numpy as np
import pytest

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


class Model:
    def __init__(self, size):
        self.free = [(0, size)]  # sorted (offset, length)

    def alloc(self, n):
        n = (max(n, 1) + 255) // 256 * 256
        for k, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    del self.free[k]
                else:
                    self.free[k] = (off + n, ln - n)
                return off
        return None

    def release(self, off, n):
        n = (max(n, 1) + 255) // 256 * 256
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged
        return max(ln for _, ln in self.free)


@pytest.mark.parametrize("seed", range(6))
def test_pool_matches_model(seed):
    import harmonize_amd as hz

    rng = np.random.default_rng(seed)
    size = 1 << 20
    model = Model(size)
    ops, live, expect = [], [], []
    for _ in range(400):
        alive = [k for k, (o, _) in enumerate(live) if o is not None]
        if alive and rng.random() < 0.45:
            k = int(rng.choice(alive))
            off, n = live[k]
            ops.append(-k - 1)
            expect.append(model.release(off, n))
            live[k] = (None, n)
        else:
            n = int(rng.integers(1, 1 << 15))
            off = model.alloc(n)
            ops.append(n)
            expect.append(NONE if off is None else off)
            live.append((off, n))
    got = hz.pool_test(size, ops)
    assert [int(x) for x in got] == [int(x) for x in expect]


def test_pool_rejects_bad_free():
    import harmonize_amd as hz

    with pytest.raises(hz.HarmonizeError):
        hz.pool_test(4096, [100, -1, -1])  # the same allocation freed twice
