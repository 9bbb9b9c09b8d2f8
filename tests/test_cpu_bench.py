"""CPU checks of the benchmark's inputs: the committed HBM-traffic profile parses (the bench reports it as
roofline.traffic), the pool-sharing rule of the served run, and the refusal of more than one GPU."""
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


def test_more_than_one_gpu_is_refused_before_any_device_work():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so, before it loads the library."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
