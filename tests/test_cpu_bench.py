"""CPU checks of the benchmark's inputs: the committed HBM-traffic profile parses (the bench reports it as
roofline.traffic), and the pool-sharing rule of the served run."""
import bench


def test_traffic_profile_parses():
    t = bench.measured_traffic()
    assert isinstance(t, int) and t > 0


class _Info:
    def __init__(self, blocks, content):
        self.blocks_count, self.bytes_in_block_content = blocks, content


class _Tree:
    def __init__(self, blocks, content=1000):
        self._i = _Info(blocks, content)

    def info(self):
        return self._i


def test_pool_shares_water_fill():
    bb = 1000 + 36
    trees = [_Tree(10), _Tree(100), _Tree(10_000)]
    shares = bench.pool_shares(trees, 1000 * bb)
    assert shares[0] == 10 and shares[1] == 100  # small scenes are fully resident
    assert shares[2] == 890  # the rest of the budget
    assert sum(s * bb for s in shares) <= 1000 * bb
    assert bench.pool_shares(trees, 10**9) == [10, 100, 10_000]
