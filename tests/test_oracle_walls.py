"""Pins of the oracle's walls step (MainLoopWalls, idpy/LBM/MultiPhase.py:
1226-1309; no GPU): reduction to the periodic step, an exact bounce-back
permutation, a stationary rest state, fluid mass conservation and an
independent numpy gather restatement of the push-with-bounce-back rule."""
import numpy as np
import pytest

from helpers import droplet_field, rel_err

G = -5.0 / 3.0


def _channel(dims, seed):
    """Solid plates at y = 0 and y = L1 - 1 plus a solid block; 1 fluid."""
    rng = np.random.default_rng(seed)
    w = np.ones(tuple(reversed(dims)), dtype=np.uint8)
    if len(dims) == 2:
        w[0, :] = 0
        w[-1, :] = 0
        w[3:6, 4:7] = 0
    else:
        w[:, 0, :] = 0
        w[:, -1, :] = 0
        w[2:4, 3:5, 2:5] = 0
    psi_w = rng.uniform(0.2, 0.8, size=w.shape)
    return w, psi_w


def _state(xi, dims, n0):
    from oracle import cpuref
    st = cpuref.State(xi, dims)
    cpuref.init_pop(st, n0.ravel(), np.zeros(st.V * st.D))
    return st


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (23, 17)),
                                        ('D3Q19', 'D3E4', (9, 8, 7))])
def test_all_fluid_walls_equal_the_periodic_step(xi, es, dims):
    from oracle import cpuref
    n0 = droplet_field(dims, seed=41)
    a, b = _state(xi, dims, n0), _state(xi, dims, n0)
    cpuref.meta_run(a, es, G, 1.0, nsteps=7)
    cpuref.meta_run(b, es, G, 1.0, nsteps=7, walls=np.ones(b.V, dtype=np.uint8))
    assert np.array_equal(a.pop, b.pop)


@pytest.mark.parametrize('xi,dims', [('D2Q9', (9, 7)), ('D3Q19', (7, 6, 5))])
def test_bounce_back_is_an_exact_reflection(xi, dims):
    """Pure streaming (G = 0, omega ~ 0): a spike moving into a solid site
    comes back at its own site in the opposite direction after one step; a
    spike moving along the wall streams normally."""
    from oracle import cpuref
    from oracle.stencils import XI, XI_W
    xis = XI[xi]
    Q, D, V = len(xis), len(dims), int(np.prod(dims))
    walls = np.ones(tuple(reversed(dims)), dtype=np.uint8)
    if D == 2:
        walls[0, :] = 0
    else:
        walls[:, 0, :] = 0
    flat = walls.ravel()
    st = cpuref.State(xi, dims)
    f = np.array([float(w) for w in XI_W[xi]])[:, None] * np.ones(V)
    site = 3 + dims[0] * 1  # (x=3, y=1): fluid, next to the plate at y = 0
    down = next(q for q, v in enumerate(xis) if v[1] == -1 and v[0] == 1 and (D == 2 or v[2] == 0))
    up = next(q for q, v in enumerate(xis) if list(v) == [-c for c in xis[down]])
    along = next(q for q, v in enumerate(xis) if v[0] == 1 and v[1] == 0 and (D == 2 or v[2] == 0))
    f[down, site] += 0.5
    f[along, site] += 0.25
    st.pop[:] = f.ravel()
    st.psi[:] = 0.0
    cpuref.meta_run(st, 'D2E4' if D == 2 else 'D3E4', 0.0, 1e300, nsteps=1, walls=flat)
    out = st.pop.reshape(Q, V)
    base = np.array([float(w) for w in XI_W[xi]])
    assert out[up, site] == base[up] + 0.5
    assert out[along, site + 1] == base[along] + 0.25
    fl = flat == 1
    assert np.isclose(out[:, fl].sum(), f[:, fl].sum(), rtol=0, atol=1e-13)


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (16, 12)),
                                        ('D3Q19', 'D3E4', (8, 8, 6))])
def test_rest_state_with_matching_wall_psi_is_stationary(xi, es, dims):
    """Uniform density, psi_w = psi(n): no force, and bounce-back maps the
    rest equilibrium onto itself (w_q = w_opp(q))."""
    from oracle import cpuref
    walls, _ = _channel(dims, 1)
    n0 = np.full(walls.shape, 1.3)
    st = _state(xi, dims, n0)
    st.psi[:] = 1.0 - np.exp(-1.3)
    f0 = st.pop.copy().reshape(st.Q, st.V)
    cpuref.meta_run(st, es, G, 1.0, nsteps=40, walls=walls.ravel())
    fl = walls.ravel() == 1
    assert np.abs(st.pop.reshape(st.Q, st.V)[:, fl] - f0[:, fl]).max() < 1e-14


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (23, 17)),
                                        ('D3Q19', 'D3E4', (9, 8, 7))])
def test_walls_conserve_fluid_mass(xi, es, dims):
    from oracle import cpuref
    walls, psi_w = _channel(dims, 2)
    n0 = droplet_field(dims, seed=43)
    st = _state(xi, dims, n0)
    st.psi[:] = psi_w.ravel()
    fl = walls.ravel() == 1
    m0 = st.pop.reshape(st.Q, st.V)[:, fl].sum()
    cpuref.meta_run(st, es, G, 1.0, nsteps=150, walls=walls.ravel())
    m1 = st.pop.reshape(st.Q, st.V)[:, fl].sum()
    assert abs(m1 - m0) < 1e-13 * m0


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (23, 17)), ('D2Q9', 'D2E8', (19, 16)),
                                        ('D3Q19', 'D3E4', (9, 8, 7))])
def test_walls_oracle_matches_numpy_gather(xi, es, dims):
    from oracle import cpuref, np_ref
    walls, psi_w = _channel(dims, 3)
    n0 = droplet_field(dims, seed=44)
    st = _state(xi, dims, n0)
    st.psi[:] = psi_w.ravel()
    shape = (st.Q,) + walls.shape
    f = st.pop.reshape(shape).copy()
    old = st.pop_swap.reshape(shape).copy()
    for k in range(4):
        f_new, n, psi = np_ref.step_walls(f, xi, es, G, 1.0, walls, psi_w, old)
        old, f = f, f_new
    cpuref.meta_run(st, es, G, 1.0, nsteps=4, walls=walls.ravel())
    fl = walls == 1
    got = st.pop.reshape(shape)[:, fl]
    assert rel_err(got.ravel(), f[:, fl].ravel()) < 1e-13
    assert rel_err(st.n.reshape(walls.shape)[fl], n[fl]) < 1e-13
    assert np.array_equal(st.psi.reshape(walls.shape)[~fl], psi_w[~fl])


# ------------------------------------- legacy-kernel wall loops with gravity
@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (23, 17)),
                                        ('D3Q19', 'D3E4', (9, 8, 7))])
def test_legacy_walls_all_fluid_no_gravity_equal_legacy_loop(xi, es, dims):
    from oracle import cpuref
    n0 = droplet_field(dims, seed=51)
    a, b = _state(xi, dims, n0), _state(xi, dims, n0)
    cpuref.legacy_run(a, es, G, 1.0, nsteps=6)
    cpuref.legacy_walls_run(b, es, G, 1.0, np.ones(b.V, dtype=np.uint8), 0.0, nsteps=6)
    assert np.array_equal(a.pop, b.pop)
    assert np.array_equal(a.u, b.u)


@pytest.mark.parametrize('xi,dims,tau', [('D2Q9', (12, 10), 1.0), ('D2Q9', (12, 10), 0.7),
                                         ('D3Q19', (6, 5, 4), 0.8)])
def test_gravity_adds_exactly_g_momentum_per_site_and_step(xi, dims, tau):
    """G = 0, periodic, uniform density at rest: with Guo forcing the total
    momentum grows by the force exactly, sum_q xi f* = J + F (for any tau),
    so after N steps J_y = N V g and J_x = 0."""
    from oracle import cpuref
    from oracle.stencils import XI
    st = _state(xi, dims, np.full(tuple(reversed(dims)), 1.2))
    g, N = 3e-5, 25
    cpuref.legacy_walls_run(st, 'D2E4' if len(dims) == 2 else 'D3E4', 0.0, tau,
                            np.ones(st.V, dtype=np.uint8), g, nsteps=N)
    f = st.pop.reshape(st.Q, st.V)
    J = [sum(v[d] * f[q].sum() for q, v in enumerate(XI[xi])) for d in range(len(dims))]
    assert abs(J[1] - N * st.V * g) < 1e-12 * N * st.V * g
    assert abs(J[0]) < 1e-15


@pytest.mark.parametrize('xi,es,dims,g', [('D2Q9', 'D2E4', (23, 17), 0.0),
                                          ('D2Q9', 'D2E4', (23, 17), 2e-4),
                                          ('D3Q19', 'D3E4', (9, 8, 7), 1e-4)])
def test_legacy_walls_fluid_sites_match_push_bounce_back(xi, es, dims, g):
    """Pull streaming + K_HalfWayBounceBack and the push stream with half-way
    bounce-back give the same fluid-site populations; with gravity, against
    the numpy gather with the body force."""
    from oracle import cpuref, np_ref
    walls, psi_w = _channel(dims, 7)
    n0 = droplet_field(dims, seed=52)
    st = _state(xi, dims, n0)
    st.psi[:] = psi_w.ravel()
    shape = (st.Q,) + walls.shape
    f = st.pop.reshape(shape).copy()
    old = st.pop_swap.reshape(shape).copy()
    for k in range(5):
        f_new, n, psi = np_ref.step_walls(f, xi, es, G, 1.0, walls, psi_w, old, gy=g)
        old, f = f, f_new
    cpuref.legacy_walls_run(st, es, G, 1.0, walls.ravel(), g, nsteps=5)
    fl = walls == 1
    assert rel_err(st.pop.reshape(shape)[:, fl].ravel(), f[:, fl].ravel()) < 1e-13
    # the legacy loop zeroes the solid sites' populations after bouncing them
    assert not st.pop.reshape(shape)[:, ~fl].any() or np.abs(
        st.pop.reshape(shape)[0, ~fl]).max() == 0.0
    if g == 0.0:
        mt = _state(xi, dims, n0)
        mt.psi[:] = psi_w.ravel()
        cpuref.meta_run(mt, es, G, 1.0, nsteps=5, walls=walls.ravel())
        assert rel_err(st.pop.reshape(shape)[:, fl].ravel(),
                       mt.pop.reshape(shape)[:, fl].ravel()) < 1e-13


# ------------------------------------------------ Shan-Chen forcing loops
@pytest.mark.parametrize('xi,es,dims,tau', [('D2Q9', 'D2E4', (17, 13), 1.0),
                                            ('D2Q9', 'D2E8', (16, 15), 0.8),
                                            ('D3Q19', 'D3E4', (8, 7, 6), 1.0)])
def test_sc_forcing_matches_numpy_and_its_velocity_identity(xi, es, dims, tau):
    """MainLoopSimpleSC: the oracle against the numpy restatement, and the
    identity the scheme implies for the written velocity: sum xi f* / n =
    u + F / n, so u_written = u + F / (2n) (F from the numpy force on the
    oracle's own psi)."""
    from oracle import cpuref, np_ref
    from oracle.stencils import XI
    n0 = droplet_field(dims, seed=61)
    st = _state(xi, dims, n0)
    shape = (st.Q,) + tuple(reversed(dims))
    f = st.pop.reshape(shape).copy()
    for k in range(4):
        f, n, psi, uw = np_ref.step_sc(f, xi, es, G, tau)
    cpuref.legacy_walls_run(st, es, G, tau, None, forcing='sc', nsteps=4)
    assert rel_err(st.pop, f.ravel()) < 1e-13
    D, V = st.D, st.V
    assert rel_err(st.u, np.concatenate([c.ravel() for c in uw])) < 1e-12
    # one more oracle step from a copy: check the velocity identity
    prev = st.pop.reshape(shape).copy()
    n_prev = prev.sum(axis=0)
    F = np_ref.sc_force(np_ref.psi_fn(n_prev), es, G)
    j = [sum(XI[xi][q][d] * prev[q] for q in range(st.Q)) for d in range(D)]
    cpuref.legacy_walls_run(st, es, G, tau, None, forcing='sc', nsteps=1)
    want = np.concatenate([(j[d] / n_prev + 0.5 * F[d] / n_prev).ravel() for d in range(D)])
    assert np.abs(st.u - want).max() < 1e-15 * 1e2


@pytest.mark.parametrize('xi,es,dims', [('D2Q9', 'D2E4', (23, 17)),
                                        ('D3Q19', 'D3E4', (9, 8, 7))])
def test_sc_forcing_walls_all_fluid_and_mass(xi, es, dims):
    """MainLoopSimpleWallsSC with all-fluid flags equals MainLoopSimpleSC
    bit-exactly; with plates and a block it conserves the fluid mass."""
    from oracle import cpuref
    n0 = droplet_field(dims, seed=62)
    a, b = _state(xi, dims, n0), _state(xi, dims, n0)
    cpuref.legacy_walls_run(a, es, G, 1.0, None, forcing='sc', nsteps=5)
    cpuref.legacy_walls_run(b, es, G, 1.0, np.ones(b.V, dtype=np.uint8), forcing='sc',
                            nsteps=5)
    assert np.array_equal(a.pop, b.pop)
    walls, psi_w = _channel(dims, 8)
    c = _state(xi, dims, n0)
    c.psi[:] = psi_w.ravel()
    fl = walls.ravel() == 1
    m0 = c.pop.reshape(c.Q, c.V)[:, fl].sum()
    cpuref.legacy_walls_run(c, es, G, 1.0, walls.ravel(), forcing='sc', nsteps=120)
    m1 = c.pop.reshape(c.Q, c.V)[:, fl].sum()
    assert abs(m1 - m0) < 1e-13 * m0
