"""Host-side checks of the ensemble and request generators (no GPU)."""
import numpy as np

from pfamd import capi
from pfamd.ensemble import NOISE_SCALE, ensemble_problems
from pfamd.serve import instance_bytes, request_stream


def test_ensemble_configuration():
    a, b = ensemble_problems(64, 7, 5000), ensemble_problems(64, 7, 5000)
    assert [(p.deck.Nx, p.deck.Ny, p.deck.Nz, p.n_seeds, p.seed_rng) for p in a] == \
           [(p.deck.Nx, p.deck.Ny, p.deck.Nz, p.n_seeds, p.seed_rng) for p in b]
    for p in a:
        d = p.deck
        assert 32 <= d.Nx <= 160 and 32 <= d.Ny <= 160 and max(d.Nx, d.Ny) <= d.Nz <= 160
        assert 5.0 <= d.taxa_resfr <= 50.0 and 1 <= p.n_seeds <= 8
        assert 0.0 <= d.AAA * NOISE_SCALE <= 0.02
        assert p.precision == capi.PF_FP32_MIXED and p.steps == 5000 and d.tipo_inicial == 6
    assert len({(p.deck.Nx, p.deck.Ny, p.deck.Nz) for p in a}) > 50  # uneven sizes


def test_effective_noise_amplitude():
    """The deck amplitude is rescaled by entrada.f90:179-185 to the drawn
    effective amplitude."""
    p = ensemble_problems(4, 7, 10)[1]
    s, _ = p.setup()
    assert abs(s.p.AAA - p.deck.AAA * NOISE_SCALE) <= 1e-15 * abs(s.p.AAA)


def test_request_stream():
    rs = request_stream(n=4000, rate=2.0, seed=11)
    t = np.array([r.arrival for r in rs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.5) < 0.03 and abs(gaps.std() - 0.5) < 0.04  # exponential
    prec = np.array([r.problem.precision for r in rs])
    assert abs(np.mean(prec == capi.PF_FP64) - 0.5) < 0.03
    steps = np.array([r.problem.steps for r in rs])
    assert steps.min() >= 1000 and steps.max() <= 20000
    e = np.array([(r.problem.deck.Nx, r.problem.deck.Ny) for r in rs])
    assert e.min() >= 32 and e.max() <= 256
    # log-uniform: the median edge is the geometric mean of the bounds
    assert abs(np.median(e[:, 0]) - np.sqrt(32 * 256)) < 6
    assert all(1 <= r.problem.n_seeds <= 16 for r in rs)
    assert [r.arrival for r in request_stream(n=10, seed=11)] == list(t[:10])


def test_instance_bytes():
    p = ensemble_problems(1, 7, 1)[0]
    assert instance_bytes(p) == 6 * p.cells * 4
