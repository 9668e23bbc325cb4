"""Initial conditions: the King model sampler against tabulated King (1966)
concentrations and the virial theorem."""
import numpy as np
import pytest

from gasoline_amd import ics


@pytest.mark.parametrize("W0,c", [(3.0, 0.67), (6.0, 1.25), (9.0, 2.12)])
def test_king_concentration(W0, c):
    # c = log10(r_t / r_0) of King's models (Binney & Tremaine, fig. 4.9)
    rr, W, m = ics._king_profile(W0)
    assert abs(np.log10(rr[-1]) - c) < 0.02
    assert abs(W[-1]) < 1e-6 and W[0] == W0


def test_king_is_in_virial_equilibrium():
    d = ics.king(3000, 6.0, seed=2)
    m = d["m"]
    r = d["r"].astype(np.float64)
    K = 0.5 * np.sum(m * np.sum(d["v"].astype(np.float64) ** 2, 1))
    dx = r[:, None, :] - r[None, :, :]
    dd = np.sqrt((dx ** 2).sum(-1))
    np.fill_diagonal(dd, np.inf)
    U = -0.5 * np.sum(m[:, None] * m[None, :] / dd)
    assert abs(2 * K / abs(U) - 1) < 0.05
    assert abs(m.sum() - 1) < 1e-12


def test_king_ensemble_offsets():
    e = ics.king_ensemble(6, seed=1, nmin=100, nmax=400)
    assert e["off"][0] == 0 and e["off"][-1] == len(e["m"])
    assert np.all(np.diff(e["off"]) >= 100) and np.all(np.diff(e["off"]) <= 400)
    assert np.all((e["W0"] >= 3) & (e["W0"] <= 9))
