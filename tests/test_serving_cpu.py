"""Host logic of the reactor server that needs no GPU: admission order across pools."""
from types import SimpleNamespace


class FakePool:
    def __init__(self, free):
        self.free = free
        self.admitted = []

    def fits(self, r):
        return r.n <= self.free

    def admit(self, r):
        self.free -= r.n
        self.admitted.append(r.rid)


def _req(rid, mech, n):
    return SimpleNamespace(rid=rid, mech=mech, n=n)


def test_fifo_is_per_pool(pkg):
    """A request too large for its pool blocks later requests of that pool only."""
    pools = {"a": FakePool(10), "b": FakePool(10)}
    waiting = [_req(0, "a", 4), _req(1, "a", 20), _req(2, "b", 5), _req(3, "a", 1), _req(4, "b", 5),
               _req(5, "b", 1)]
    still = pkg.serving.admit_fifo_per_pool(waiting, pools)
    assert pools["a"].admitted == [0]            # 3 waits behind 1 although it would fit
    assert pools["b"].admitted == [2, 4]         # not held back by pool a's blocked request
    assert [r.rid for r in still] == [1, 3, 5]   # 5 waits behind nothing of a but did not fit b


def test_admission_order_is_arrival_order_within_a_pool(pkg):
    pools = {"a": FakePool(3)}
    waiting = [_req(i, "a", 1) for i in range(5)]
    still = pkg.serving.admit_fifo_per_pool(waiting, pools)
    assert pools["a"].admitted == [0, 1, 2]
    assert [r.rid for r in still] == [3, 4]
