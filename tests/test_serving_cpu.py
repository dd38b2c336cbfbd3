"""Admission logic of the continuous-batching server (guacho_amd.serving)
on the CPU, with a stand-in batch object that counts steps instead of
running kernels: a request larger than the whole budget is rejected instead
of blocking the queue, a full box table leaves requests waiting instead of
aborting the serve, and small requests are admitted past a large one that
does not fit yet."""
import numpy as np

from guacho_amd import _capi, serving


class FakeBatch:
    """Boxes of `cells * 10` bytes; every advance() takes one step of each
    box not yet done, visible from the next query(); a box is done after
    its max_steps."""

    def __init__(self, precision, max_boxes, steps_per_launch, memory_budget):
        self.budget = memory_budget
        self.used = 0
        self.max_boxes = max_boxes
        self.boxes = {}
        self.next_id = 0
        self.log = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    def box_bytes(self, p):
        return p.nx * p.ny * p.nz * 10

    def free_bytes(self):
        return self.budget - self.used

    def add(self, p, u0, cfl, t_end, max_steps):
        if len(self.boxes) >= self.max_boxes:
            raise _capi.GmError(_capi.GM_ERR_BUSY, "box table full")
        need = self.box_bytes(p)
        if need > self.free_bytes():
            raise _capi.GmError(_capi.GM_ERR_BUSY, "memory budget exhausted")
        bid = self.next_id
        self.next_id += 1
        self.boxes[bid] = [need, 0, max_steps]
        self.used += need
        self.log.append(("add", p.nx))
        return bid

    def advance(self, n):
        # like the real batch: the launch runs later (here: at the next
        # query), boxes already completed are not stepped, and clock()
        # reports the state as of the last query
        self.pending = [bid for bid, b in self.boxes.items() if b[1] < b[2]]

    def sync(self):
        self.query()

    def query(self):
        for bid in getattr(self, "pending", []):
            if bid in self.boxes:
                self.boxes[bid][1] += 1
        self.pending = []
        return True

    def clock(self, bid):
        need, steps, mx = self.boxes[bid]
        return 0.0, steps, steps >= mx

    def download_u(self, bid):
        return np.zeros(1)

    def remove(self, bid):
        self.used -= self.boxes.pop(bid)[0]


def _req(rid, n, steps, arrival=0.0):
    p = _capi.make_params(n, 4, 4, 0.1, 0.1, 0.1)
    r = serving.Request(rid, "box", p, np.zeros(1), max_steps=steps)
    r.arrival = arrival
    return r


def _serve(reqs, budget, max_boxes=8):
    made = {}

    def factory(**kw):
        kw["memory_budget"] = budget
        made["b"] = FakeBatch(**kw)
        return made["b"]
    stats = serving.serve(reqs, steps_per_launch=1, max_boxes=max_boxes, batch_factory=factory)
    return stats, made["b"]


def test_oversized_request_is_rejected_not_blocking():
    # budget 1000 bytes: the 64x4x4 box needs 10240 and can never run
    reqs = [_req(0, 64, 3), _req(1, 4, 2), _req(2, 4, 2)]
    stats, B = _serve(reqs, budget=1000)
    assert reqs[0].rejected and not reqs[1].rejected and not reqs[2].rejected
    assert stats["rejected"] == 1 and stats["requests"] == 2
    assert reqs[1].steps == 2 and reqs[2].steps == 2
    assert B.used == 0 and not B.boxes


def test_full_box_table_leaves_requests_waiting():
    reqs = [_req(i, 4, 2 + i % 3) for i in range(7)]
    stats, B = _serve(reqs, budget=10 ** 9, max_boxes=2)
    assert stats["requests"] == 7 and stats["rejected"] == 0
    assert all(r.steps == 2 + r.rid % 3 for r in reqs)
    assert all(r.finished >= r.admitted >= 0 for r in reqs)


def test_small_requests_pass_a_blocked_large_one():
    # budget holds the large box (8x4x4 = 1280 B) alone, or the small boxes
    # (4x4x4 = 640 B) two at a time; the large one waits for the first
    # small box to finish, and the small ones behind it are not held back
    reqs = [_req(0, 4, 5), _req(1, 8, 1), _req(2, 4, 1), _req(3, 4, 1)]
    stats, B = _serve(reqs, budget=1300)
    order = [n for kind, n in B.log if kind == "add"]
    assert order[:2] == [4, 4], order          # request 2 admitted past the large request 1
    assert sorted(order) == [4, 4, 4, 8]
    assert stats["requests"] == 4
    assert all(r.steps >= 1 for r in reqs)
