"""GpuBatchSet's bookkeeping (which batch a problem joins, which batches run,
ids, removal, the memory budget) over stand-in batches. CPU only: the device
batches themselves are tested in test_gpu_batch.py."""
import pytest

import mocc_amd.sweeper as sw


class FakeBatch:
    """A batch whose problem i finishes after its own number of outer
    iterations; poll reports each finish once, as mocgpu_batch_poll does."""
    runs = []  # (batch, n_outer) of every run, in order

    def __init__(self, budget_bytes: int = 0, use_graph: bool = True):
        self.members = {}  # id -> [iterations left, bytes, reported]
        self.next = 0

    def add(self, s, k_tol, psi_tol, max_iter, min_iter=0):
        cid = self.next
        self.next += 1
        self.members[cid] = [s.iters, s.state_bytes(), False]
        return cid

    def run(self, n_outer=1):
        FakeBatch.runs.append((self, n_outer))
        for m in self.members.values():
            m[0] = max(0, m[0] - n_outer)

    def poll(self, cap=4096):
        out = []
        for cid, m in self.members.items():
            if m[0] == 0 and not m[2]:
                m[2] = True
                out.append((cid, f"state-{cid}"))
        return out

    def remove(self, cid):
        del self.members[cid]

    def size(self):
        return len(self.members)

    def bytes_used(self):
        return sum(m[1] for m in self.members.values())

    def close(self):
        self.members.clear()


class FakeSweeper:
    def __init__(self, iters, nbytes=100):
        self.iters, self.nbytes = iters, nbytes

    def state_bytes(self):
        return self.nbytes


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(sw, "GpuBatch", FakeBatch)
    FakeBatch.runs = []
    return FakeBatch


def test_members_spread_and_ids(fake):
    bs = sw.GpuBatchSet(2)
    ids = [bs.add(FakeSweeper(3), 1e-7, 1e-6, 100) for _ in range(5)]
    assert ids == [0, 1, 2, 3, 4]
    assert [b.size() for b in bs.batches] == [3, 2]
    assert bs.size() == 5 and bs.bytes_used() == 500


def test_only_batches_with_unfinished_members_run(fake):
    bs = sw.GpuBatchSet(2)
    quick = bs.add(FakeSweeper(1), 0, 0, 100)   # batch 0
    slow = bs.add(FakeSweeper(4), 0, 0, 100)    # batch 1
    bs.run(1)
    assert [b for b, _ in fake.runs] == bs.batches
    done = dict(bs.poll())
    assert set(done) == {quick}
    fake.runs.clear()
    bs.run(1)
    assert [b for b, _ in fake.runs] == [bs.batches[1]]  # batch 0 has only a finished member
    assert bs.poll() == []
    bs.remove(quick)
    assert bs.size() == 1
    bs.run(2)
    assert dict(bs.poll()).keys() == {slow}
    fake.runs.clear()
    bs.run(1)
    assert fake.runs == []  # nothing unfinished


def test_solve_all_and_compaction(fake):
    bs = sw.GpuBatchSet(3)
    want = {bs.add(FakeSweeper(n), 0, 0, 100): n for n in (1, 2, 3, 4, 5, 6, 7)}
    res = bs.solve_all(chunk=1, compact=0.25)
    assert set(res) == set(want)
    # finished problems were taken out on the way; the last ones stay members
    assert 0 < bs.size() < len(want)
    for sid in list(res):
        if sid in bs._where:
            bs.remove(sid)
    assert bs.size() == 0 and all(n == 0 for n in bs._live)


def test_budget_is_for_the_whole_set(fake):
    bs = sw.GpuBatchSet(2, budget_bytes=250)
    bs.add(FakeSweeper(1, 100), 0, 0, 10)
    bs.add(FakeSweeper(1, 100), 0, 0, 10)
    with pytest.raises(sw.CapacityError):
        bs.add(FakeSweeper(1, 100), 0, 0, 10)
    assert bs.size() == 2
    with pytest.raises(ValueError):
        sw.GpuBatchSet(0)
