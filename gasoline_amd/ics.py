"""Seeded synthetic initial conditions for the force-step workloads.

Positions are FP32-representable (gasoline's Tipsy reader stores FP32,
pkd.c:482-519), held in FP64 arrays.  The periodic box draws coordinates on a
2^-24 lattice so that every FP64 separation and squared distance is exact.
"""
import numpy as np


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def plummer(n, seed=1, mfrac_cut=0.999):
    """Plummer sphere, G = M = a = 1, equal masses.

    Radii from the inverse cumulative mass profile (Aarseth, Henon & Wielen
    1974) with the mass fraction capped at mfrac_cut; isotropic directions;
    speeds by rejection from g(q) = q^2 (1 - q^2)^(7/2) times the local escape
    speed sqrt(2) (1 + r^2)^(-1/4).
    Returns (r[n,3], v[n,3], m[n]).
    """
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(0.0, mfrac_cut, n)
    x1 = np.maximum(x1, 1e-10)
    rad = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    dirs = _iso(rng, n)
    pos = rad[:, None] * dirs
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0, 1, todo.size)
        yy = rng.uniform(0, 0.1, todo.size)
        ok = yy < qq * qq * (1 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0) * (1.0 + rad * rad) ** (-0.25)
    vel = (q * vesc)[:, None] * _iso(rng, n)
    m = np.full(n, 1.0 / n)
    return _f32(pos), _f32(vel), m


def _iso(rng, n):
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)


def sph_box(n, seed=2, sigma_v=0.1, u=1.0):
    """Adiabatic periodic box: n gas particles uniform in [0,1)^3 on a 2^-24
    lattice, equal mass 1/n, u = 1, Gaussian velocities (sigma 0.1 per axis).
    Returns dict of r, v, m, u."""
    rng = np.random.default_rng(seed)
    pos = rng.integers(0, 1 << 24, size=(n, 3)).astype(np.float64) / float(1 << 24)
    vel = _f32(rng.normal(0.0, sigma_v, size=(n, 3)))
    return dict(r=pos, v=vel, m=np.full(n, 1.0 / n), u=np.full(n, float(u)))


def evrard(n, seed=5, u=0.05):
    """Evrard collapse: cold gas sphere rho ~ 1/r inside R = 1, M = G = 1,
    u = 0.05, at rest.  Radii r = sqrt(X) (enclosed mass ~ r^2)."""
    rng = np.random.default_rng(seed)
    rad = np.sqrt(rng.uniform(0, 1, n))
    pos = rad[:, None] * _iso(rng, n)
    return dict(r=_f32(pos), v=np.zeros((n, 3)), m=np.full(n, 1.0 / n),
                u=np.full(n, float(u)))


def _king_profile(W0):
    """King (1966) model of central potential W0 = Psi(0)/sigma^2: radius in
    units of the King radius r0 = sqrt(9 sigma^2 / (4 pi G rho0)), W(r), and
    the enclosed mass in units of 4 pi rho0 r0^3, out to the tidal radius."""
    from scipy.integrate import solve_ivp
    from scipy.special import erf

    def rho(W):
        W = np.maximum(W, 0.0)
        return np.exp(W) * erf(np.sqrt(W)) - np.sqrt(4 * W / np.pi) * (1 + 2 * W / 3)

    rho0 = rho(np.array(W0))

    def rhs(r, y):
        W, dW, m = y
        d = rho(np.array(W)) / rho0
        return [dW, -9.0 * d - (2.0 / r) * dW, d * r * r]

    def tidal(r, y):
        return y[0]
    tidal.terminal = True
    r0 = 1e-6
    y0 = [W0 - 1.5 * r0 * r0, -3.0 * r0, r0 ** 3 / 3]   # series: W ~ W0 - 3/2 r^2
    sol = solve_ivp(rhs, (r0, 1e4), y0, events=tidal, max_step=0.05, rtol=1e-9, atol=1e-12,
                    dense_output=True)
    rt = sol.t_events[0][0]
    rr = np.concatenate([[0.0], np.linspace(r0, rt, 4000)])
    Y = sol.sol(np.clip(rr, r0, rt))
    W = np.maximum(Y[0], 0.0)
    W[0] = W0
    m = np.maximum(Y[2], 0.0)
    m[0] = 0.0
    return rr, W, m


def king(n, W0, seed=0):
    """One King model: n equal-mass particles, total mass M = 1, G = 1,
    lengths in King radii (r0 = 1).  Radii by the inverse cumulative mass,
    isotropic directions, speeds by rejection from v^2 (exp(W - v^2/2) - 1)
    in units of sigma = sqrt(1 / (9 mu)), mu the dimensionless total mass."""
    rng = np.random.default_rng(seed)
    rr, W, m = _king_profile(float(W0))
    mu = m[-1]
    rad = np.interp(rng.uniform(0, mu, n), m, rr)
    Wr = np.interp(rad, rr, W)
    v = np.empty(n)
    todo = np.arange(n)
    while len(todo):      # rejection, vectorised over the particles still open
        Wt = Wr[todo]
        vmax = np.sqrt(2 * Wt)
        fmax = vmax * vmax * (np.exp(Wt) - 1) + 1e-300
        x = rng.uniform(0, 1, len(todo)) * vmax
        ok = rng.uniform(0, 1, len(todo)) * fmax <= x * x * (np.exp(Wt - 0.5 * x * x) - 1)
        v[todo[ok]] = x[ok]
        todo = todo[~ok]
    sigma = np.sqrt(1.0 / (9.0 * mu))
    pos = rad[:, None] * _iso(rng, n)
    vel = (v * sigma)[:, None] * _iso(rng, n)
    return dict(r=_f32(pos), v=_f32(vel), m=np.full(n, 1.0 / n))


def king_ensemble(nsys, seed=0, nmin=1024, nmax=65536, W0=(3.0, 9.0)):
    """nsys King models with W0 uniform in [W0[0], W0[1]] and N log-uniform in
    [nmin, nmax], stored one after the other; returns the concatenated
    particles and the offsets of the systems."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys)).astype(int)
    ws = rng.uniform(W0[0], W0[1], nsys)
    parts = [king(int(ns[k]), ws[k], seed=seed * 1000 + k) for k in range(nsys)]
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int32)
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("r", "v", "m")}
    cat["off"], cat["W0"] = off, ws
    return cat


# Galactic units of the disk-galaxy workload: lengths in kpc, velocities in
# km/s, G = 1; the mass unit is then 1 / G[kpc (km/s)^2 / Msun].
G_KPC_KMS2_MSUN = 4.30091e-6
MASS_UNIT_MSUN = 1.0 / G_KPC_KMS2_MSUN
K_BOLTZ, M_PROTON = 1.380649e-23, 1.67262192e-27


def _hernquist_sigma2(r, M, a):
    """Isotropic radial velocity dispersion of a Hernquist sphere (Hernquist
    1990, eq. 10), G = 1."""
    x = r / a
    t = 12.0 * x * (1 + x) ** 3 * np.log((1 + x) / x) \
        - x / (1 + x) * (25 + 52 * x + 42 * x * x + 12 * x ** 3)
    return np.maximum(M / (12.0 * a) * t, 0.0)


def disk_galaxy(n_dm=1 << 22, n_gas=1 << 20, seed=11, M_halo=1e12, a_halo=30.0, r_cut=10.0,
                R_d=3.0, z_0=0.3, Q=1.5, T=1e4, mu=1.22, gamma=5.0 / 3.0,
                eps_dm=0.1, eps_gas=0.05):
    """Isolated disk galaxy in galactic units (kpc, km/s, G = 1):

    * Hernquist halo, M = M_halo Msun, scale a_halo kpc, truncated at
      r_cut * a_halo (its particles carry the truncated mass); radii from the
      inverse cumulative mass, isotropic Gaussian velocities with the Jeans
      dispersion of the untruncated profile.
    * Exponential gas disk, scale length R_d, sech^2(z / z_0) layers, cut at
      10 R_d; cold circular rotation in the halo + disk potential (spherical
      approximation of the disk's enclosed mass); temperature T with mean
      molecular weight mu, so u = kT / ((gamma - 1) mu m_p); the disk mass is
      set so that the minimum over R of the gas Toomre parameter
      Q = c_s kappa / (pi G Sigma) (halo epicyclic frequency) equals Q.

    Returns dict(r, v, m, u, soft, gas) with gas a boolean mask (the first
    n_dm particles are the halo) and the derived disk mass in M_disk (Msun).
    """
    rng = np.random.default_rng(seed)
    Mh = M_halo / MASS_UNIT_MSUN
    f_cut = (r_cut / (1.0 + r_cut)) ** 2
    X = rng.uniform(0.0, f_cut, n_dm)
    sx = np.sqrt(X)
    rad = a_halo * sx / (1.0 - sx)
    r_dm = rad[:, None] * _iso(rng, n_dm)
    sig = np.sqrt(_hernquist_sigma2(np.maximum(rad, 1e-6 * a_halo), Mh, a_halo))
    v_dm = sig[:, None] * rng.normal(size=(n_dm, 3))
    # gas: u and sound speed from T (km/s)^2
    kT_mu = K_BOLTZ * T / (mu * M_PROTON) / 1e6
    u = kT_mu / (gamma - 1.0)
    cs = np.sqrt(gamma * kT_mu)
    # disk mass from min Q over R (halo epicyclic frequency)
    R = np.linspace(0.05 * R_d, 10 * R_d, 4000)
    vc2h = Mh * R / (R + a_halo) ** 2
    om2 = vc2h / R ** 2
    kap = np.sqrt(np.maximum(R * np.gradient(om2, R) + 4 * om2, 0.0))
    # Q = cs kap / (pi Sigma), Sigma = Md / (2 pi R_d^2) e^{-R/R_d}
    Md = np.min(cs * kap * 2.0 * R_d ** 2 * np.exp(R / R_d) / Q)
    Rg = np.empty(n_gas)
    todo = np.arange(n_gas)
    while todo.size:
        rr = -R_d * np.log(rng.uniform(size=todo.size) * rng.uniform(size=todo.size))
        ok = rr < 10 * R_d
        Rg[todo[ok]] = rr[ok]
        todo = todo[~ok]
    ph = rng.uniform(0, 2 * np.pi, n_gas)
    z = z_0 * np.arctanh(np.clip(rng.uniform(-1, 1, n_gas), -1 + 1e-12, 1 - 1e-12))
    r_g = np.stack([Rg * np.cos(ph), Rg * np.sin(ph), z], axis=1)
    m_enc = Mh * Rg ** 2 / (Rg + a_halo) ** 2 + Md * (1 - (1 + Rg / R_d) * np.exp(-Rg / R_d))
    vc = np.sqrt(m_enc / np.maximum(Rg, 1e-6))
    v_g = np.stack([-vc * np.sin(ph), vc * np.cos(ph), np.zeros(n_gas)], axis=1)
    r = np.concatenate([r_dm, r_g])
    v = np.concatenate([v_dm, v_g])
    m = np.concatenate([np.full(n_dm, Mh * f_cut / n_dm), np.full(n_gas, Md / n_gas)])
    gas = np.zeros(n_dm + n_gas, bool)
    gas[n_dm:] = True
    soft = np.where(gas, eps_gas, eps_dm)
    return dict(r=_f32(r), v=_f32(v), m=m, u=np.where(gas, u, 0.0), soft=soft, gas=gas,
                M_disk=Md * MASS_UNIT_MSUN, cs=cs, units="kpc, km/s, G=1, mass unit %.6g Msun"
                % MASS_UNIT_MSUN)
