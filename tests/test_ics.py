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


def test_hernquist_dispersion_closed_form():
    # at r = a (Hernquist 1990 eq. 10): sigma^2 = GM/(12a) (96 ln 2 - 65.5)
    s2 = ics._hernquist_sigma2(np.array([30.0]), 1.0, 30.0)[0]
    assert abs(s2 - (96 * np.log(2) - 65.5) / 360.0) < 1e-12
    # the isotropic Jeans equation: rho sigma^2 (r) = int_r^inf rho G M / r'^2
    from scipy.integrate import quad
    rho = lambda r: 1.0 / (2 * np.pi * r * (r + 1.0) ** 3)
    for r in (0.1, 0.5, 2.0, 8.0):
        I = quad(lambda x: rho(x) * x * x / (x + 1.0) ** 2 / x ** 2, r, np.inf,
                 epsabs=0, epsrel=1e-12, limit=200)[0]
        assert abs(ics._hernquist_sigma2(np.array([r]), 1.0, 1.0)[0] * rho(r) / I - 1) < 1e-6


def test_disk_galaxy_profiles():
    d = ics.disk_galaxy(1 << 17, 1 << 16, seed=3)
    g = d["gas"]
    assert g.sum() == 1 << 16 and (~g).sum() == 1 << 17
    R = np.hypot(d["r"][g, 0], d["r"][g, 1])
    # exponential disk cut at 10 R_d: mean R = R_d (2 - 100 e^-10/(1 - 11 e^-10))
    mean_R = 3.0 * (2 - 110 * np.exp(-10) / (1 - 11 * np.exp(-10)))
    assert abs(R.mean() / mean_R - 1) < 0.01
    # sech^2 layers: std z = z_0 pi / sqrt(12)
    assert abs(d["r"][g, 2].std() / (0.3 * np.pi / np.sqrt(12)) - 1) < 0.02
    # Hernquist halo cut at 10 a: the median radius encloses half the particles
    # of M(<r) = M r^2/(r+a)^2, i.e. r/(r+a) = sqrt(0.5 (10/11)^2)
    q = np.sqrt(0.5) * 10 / 11
    rh = np.linalg.norm(d["r"][~g], axis=1)
    assert abs(np.median(rh) / (30 * q / (1 - q)) - 1) < 0.02
    # halo particles carry the truncated mass; the gas has u = kT/((gamma-1) mu m_p)
    assert abs(d["m"][~g].sum() * ics.MASS_UNIT_MSUN / (1e12 * (10 / 11) ** 2) - 1) < 1e-12
    kT = 1.380649e-23 * 1e4 / (1.22 * 1.67262192e-27) / 1e6
    assert np.allclose(d["u"][g], 1.5 * kT) and np.all(d["u"][~g] == 0)
    # the disk mass puts the minimum Toomre Q at 1.5: Q(R) >= 1.5 on a grid
    Rg = np.linspace(0.3, 30, 2000)
    Mh = 1e12 / ics.MASS_UNIT_MSUN
    om2 = Mh / (Rg + 30) ** 2 / Rg
    kap = np.sqrt(np.gradient(om2, Rg) * Rg + 4 * om2)
    Sig = d["M_disk"] / ics.MASS_UNIT_MSUN / (2 * np.pi * 9) * np.exp(-Rg / 3)
    Qr = d["cs"] * kap / (np.pi * Sig)
    assert abs(Qr.min() - 1.5) < 0.01
