"""Host-side checks of the serving loop (pfamd.serve.serve_batched) against
a stand-in for the batched form: a Python batch that keeps each instance's
steps left and byte count and enforces the HBM budget as pf_batch_add does.
No GPU: this tests the loop's admission, overtaking, tick and retirement
logic, not the kernels (tests/test_gpu_serve.py runs the real batch)."""
import numpy as np
import pytest

from pfamd import capi, serve
from pfamd.deck import Deck
from pfamd.ensemble import Problem


class FakeBatch:
    """The subset of capi.Batch the serving loop calls."""

    instances = []  # every FakeBatch made (the test reads the last one)

    def __init__(self, hbm_budget=0, lib=None):
        self.budget = hbm_budget or (1 << 40)
        self.slots = {}  # slot -> [steps_left, steps_done, bytes, precision, nx, ny]
        self.runs = []  # (k, steps left of every slot before the run)
        self.adds = []
        self.peak = 0
        FakeBatch.instances.append(self)

    def set_graph_steps(self, k):
        pass

    def bytes(self):
        return sum(s[2] for s in self.slots.values()), self.budget

    def add(self, p, L, nsteps, phi, c, T):
        nbytes = 6 * p.nx * p.ny * p.nz * (8 if p.precision == capi.PF_FP64 else 4)
        if self.bytes()[0] + nbytes > self.budget:
            raise capi.PfError(capi.PF_E_NOMEM, "over the budget")
        slot = max(self.slots, default=-1) + 1
        self.slots[slot] = [nsteps, 0, nbytes, p.precision, p.nx, p.ny]
        self.adds.append(slot)
        self.peak = max(self.peak, self.bytes()[0])
        return slot

    def status(self, slot):
        s = self.slots[slot]
        return s[1], s[0]

    def run(self, k):
        self.runs.append((k, {sl: s[0] for sl, s in self.slots.items()}))
        for s in self.slots.values():
            d = min(k, s[0])
            s[0] -= d
            s[1] += d

    def synchronize(self):
        pass

    def remove(self, slot):
        del self.slots[slot]

    def size_class(self, precision, nx, ny):
        return precision * 4 + (0 if nx * ny <= 48 * 48 else 3)

    def autotune(self, reps=1, size_class=-1):
        pass

    def close(self):
        pass


class _P:  # the fields of pf_params the stand-in reads
    def __init__(self, deck, precision):
        self.nx, self.ny, self.nz, self.precision = deck.Nx, deck.Ny, deck.Nz, precision


class _Setup:
    def __init__(self, deck, precision):
        self.p, self.L = _P(deck, precision), None


def _request(rid, arrival, edge, steps, precision=capi.PF_FP32_MIXED):
    deck = Deck(Nx=edge, Ny=edge, Nz=edge, tipo_inicial=6, R_semente=4.0, ruido_tipo=1,
                taxa_resfr=10.0, n_varreduras=2, stoptime=0.0)
    return serve.Request(rid, arrival, Problem(deck=deck, precision=precision, steps=steps))


@pytest.fixture
def fake(monkeypatch):
    FakeBatch.instances.clear()
    monkeypatch.setattr(serve.capi, "Batch", FakeBatch)
    monkeypatch.setattr(Problem, "setup",
                        lambda self, lib=None: (_Setup(self.deck, self.precision),
                                                (None, None, None, None, None)))
    return FakeBatch


def _bytes(r):
    return serve.instance_bytes(r.problem)


def test_every_request_completes_once_within_the_budget(fake):
    rng = np.random.default_rng(5)
    reqs = [_request(q, 0.002 * q, int(rng.integers(8, 40)), int(rng.integers(1, 900)),
                     capi.PF_FP64 if q % 2 else capi.PF_FP32_MIXED) for q in range(40)]
    budget = 3 * max(_bytes(r) for r in reqs)
    rep = serve.serve_batched(reqs, hbm_budget=budget, tick_steps=200, lib=object())
    b = fake.instances[-1]
    assert sorted(c.rid for c in rep.completions) == list(range(40))
    assert b.peak <= budget and rep.peak_bytes <= budget
    assert not b.slots  # every instance retired
    for c in rep.completions:
        assert c.arrival <= c.admitted <= c.done
    # a tick runs at most tick_steps and, ending at the first completion,
    # never past the smallest step count left
    for k, left in b.runs:
        assert 1 <= k <= 200 and k <= min(left.values())
    assert rep.busy_ticks == len(b.runs) == len(rep.tick_log)


def test_ticks_run_full_length_without_end_at_completion(fake):
    reqs = [_request(0, 0.0, 10, 450), _request(1, 0.0, 10, 130)]
    serve.serve_batched(reqs, tick_steps=200, end_at_completion=False, lib=object())
    b = fake.instances[-1]
    # ticks of min(200, largest left): 200, 200, then the 50 left
    assert [k for k, _ in b.runs] == [200, 200, 50]


def test_overtaking_is_bounded(fake):
    """A head that does not fit waits; smaller requests behind it overtake
    it at most overtake_limit times in a row, then admission waits for the
    head."""
    big = _request(1, 0.0, 40, 50)
    small = [_request(2 + q, 0.0, 8, 5000) for q in range(6)]
    first = _request(0, 0.0, 40, 400)  # occupies the budget when the big one arrives
    budget = _bytes(first) + _bytes(small[0]) * 6  # big does not fit beside first
    rep = serve.serve_batched([first, big] + small, hbm_budget=budget, overtake_limit=2,
                              tick_steps=100, lib=object())
    assert rep.overtakes == 2
    order = [sl for sl in fake.instances[-1].adds]
    assert len(order) == 8
    done = {c.rid: c for c in rep.completions}
    # the two overtakers were admitted before the big request, the rest after it
    adm = sorted(done.values(), key=lambda c: (c.admitted, c.rid))
    assert [c.rid for c in adm][:4] == [0, 2, 3, 1]


def test_request_too_large_for_the_budget_raises(fake):
    r = _request(0, 0.0, 30, 10)
    with pytest.raises(RuntimeError, match="needs more than the HBM budget"):
        serve.serve_batched([r], hbm_budget=_bytes(r) - 1, lib=object())
