"""numpy brute force of one Shan-Chen step (TEST INFRASTRUCTURE ONLY).

An independent, vectorised restatement (np.roll over lattice and E vectors)
used only to pin the C++ oracle on small grids, following the pattern of the
reference's periodic-convolution check (idpy/test_convolution.py:92-160).
Fields are numpy arrays of shape flip(dims) (z, y, x), populations have a
leading Q axis.
"""
import numpy as np

from .stencils import E, XI, XI_W


def _shift(a, e):
    """a(x + e) on the periodic lattice; e is (e_x, e_y[, e_z])."""
    D = len(e)
    return np.roll(a, shift=tuple(-int(v) for v in e), axis=tuple(range(a.ndim - 1, a.ndim - 1 - D, -1)))


def psi_fn(n, kind='exp'):
    return {'exp': lambda n: 1.0 - np.exp(-n),
            'expinv': lambda n: np.exp(-1.0 / n),
            'linear': lambda n: n}[kind](n)


def sc_force(psi, estencil, G):
    """F_d = -G psi(x) sum_e w_e e_d psi(x+e) (IDShanChenForce.py:85-114)."""
    es, ws = E[estencil]
    D = len(es[0])
    grad = [np.zeros_like(psi) for _ in range(D)]
    for e, w in zip(es, ws):
        sp_ = _shift(psi, e)
        for d in range(D):
            if e[d]:
                grad[d] += e[d] * float(w) * sp_
    return [-G * psi * g for g in grad]


def feq_guo(xi, n, u, F):
    """Second-order Hermite equilibrium and Guo term without (1 - w/2)
    (Equilibria.py:107-138, Forcings.py:104-127)."""
    c2 = 1.0 / 3.0
    uu = sum(ud * ud for ud in u)
    Fu = sum(Fd * ud for Fd, ud in zip(F, u))
    feq, fg = [], []
    for v, w in zip(XI[xi], XI_W[xi]):
        w = float(w)
        xu = sum(v[d] * u[d] for d in range(len(u)))
        xF = sum(v[d] * F[d] for d in range(len(u)))
        feq.append(w * n * (1 + xu / c2 + xu * xu / (2 * c2 * c2) - uu / (2 * c2)))
        fg.append(w * (xF / c2 + xF * xu / (c2 * c2) - Fu / c2))
    return np.array(feq), np.array(fg)


def iso_filter_solve(f):
    """K_IsotropyFilter (idpy/LBM/LBMKernelsMeta.py:895-1080) by a dense
    solve at every site: f has shape (9, ...). Moment rows are built here
    from the lattice vectors (Hermite products h_a(xi_x) h_b(xi_y) with
    h_0 = 1, h_1 = x, h_2 = x^2 - 1/3), normalised by the density; the new
    populations solve the system whose xy, xx - yy, xxy and xyy rows take
    the reference's substituted values."""
    x = np.array([v[0] for v in XI['D2Q9']], dtype=np.float64)
    y = np.array([v[1] for v in XI['D2Q9']], dtype=np.float64)
    h2 = lambda a: a * a - 1.0 / 3.0
    one = np.ones(9)
    mrows = np.array([one, x, y, h2(x), x * y, h2(y), h2(x) * y, x * h2(y), h2(x) * h2(y)])
    arows = np.array([one, x, y, h2(x), x * y, x * x - y * y, x * x * y, x * y * y,
                      h2(x) * h2(y)])
    sh = f.shape[1:]
    ff = f.reshape(9, -1)
    m = mrows @ ff
    m0 = m[0].copy()
    m[1:] /= m0
    t = 1.0 / 3.0
    b = np.array([np.ones_like(m0), m[1], m[2], m[3], m[1] * m[2], m[1] ** 2 - m[2] ** 2,
                  2 * m[4] * m[1] + (m[3] + t) * m[2] - 2 * m[2] * m[1] ** 2,
                  2 * m[4] * m[2] + (m[5] + t) * m[1] - 2 * m[1] * m[2] ** 2, m[8]])
    out = np.linalg.solve(arows, b) * m0
    return out.reshape((9,) + sh)


def step(f, xi, estencil, G, tau, psi_kind='exp', mrt=None, iso=False):
    """One step: density/psi, force, Guo-shifted velocity, BGK (or the
    dense MRT form A f + B f_eq + Gg F_guo), push streaming; with iso
    (D2Q9) the isotropy filter between collision and stream, the step of
    MainLoopIsoFilter (idpy/LBM/MultiPhase.py:1099-1223). Returns
    (f_new, n, psi, u_shifted)."""
    omega = 1.0 / tau
    xis = XI[xi]
    D = len(xis[0])
    n = f.sum(axis=0)
    psi = psi_fn(n, psi_kind)
    F = sc_force(psi, estencil, G)
    u = [sum(xis[q][d] * f[q] for q in range(len(xis))) / n + 0.5 * F[d] / n for d in range(D)]
    feq, fg = feq_guo(xi, n, u, F)
    if mrt is None:
        fs = (1 - omega) * f + omega * feq + (1 - 0.5 * omega) * fg
    else:
        A, B, Gg = mrt
        fs = (np.tensordot(A, f, axes=1) + np.tensordot(B, feq, axes=1)
              + np.tensordot(Gg, fg, axes=1))
    if iso:
        fs = iso_filter_solve(fs)
    fn = np.empty_like(f)
    for q, v in enumerate(xis):
        fn[q] = _shift(fs[q], tuple(-c for c in v))
    return fn, n, psi, u


def step_walls(f, xi, estencil, G, tau, walls, psi_solid, f_old, psi_kind='exp', gy=0.0):
    """One step of MainLoopWalls (idpy/LBM/MultiPhase.py:1226-1309), written
    as a gather: population p of fluid site x comes from x - c_p when that
    site is fluid (push), or is its own f*_opp(p) when x - c_p is solid
    (half-way bounce-back, idpy/LBM/Collisions.py:835-850). walls: 1 fluid /
    0 solid, shape flip(dims); psi_solid: psi of the solid sites; f_old: the
    destination buffer's previous contents (kept where nothing arrives); gy:
    body force along y added to the Shan-Chen force. Returns (f_new, n, psi)."""
    omega = 1.0 / tau
    xis = XI[xi]
    D = len(xis[0])
    fluid = walls == 1
    solid = walls == 0
    n = np.where(fluid, f.sum(axis=0), 1.0)
    psi = np.where(fluid, psi_fn(n, psi_kind), psi_solid)
    F = sc_force(psi, estencil, G)
    F[1] = F[1] + gy
    u = [sum(xis[q][d] * f[q] for q in range(len(xis))) / n + 0.5 * F[d] / n for d in range(D)]
    feq, fg = feq_guo(xi, n, u, F)
    fs = (1 - omega) * f + omega * feq + (1 - 0.5 * omega) * fg
    opp = [[-c for c in v] for v in xis]
    opp = [next(k for k, w in enumerate(xis) if list(w) == o) for o in opp]
    fn = np.array(f_old, copy=True)
    for p, v in enumerate(xis):
        back = tuple(-c for c in v)
        src = _shift(fs[p], back)
        fluid_src = _shift(fluid, back)
        solid_src = _shift(solid, back)
        new = np.where(fluid_src, src, np.where(solid_src, fs[opp[p]], f_old[p]))
        fn[p] = np.where(fluid, new, f_old[p])
    return fn, np.where(fluid, n, 0.0), psi


def step_sc(f, xi, estencil, G, tau, psi_kind='exp'):
    """One periodic step with the Shan-Chen forcing (MainLoopSimpleSC,
    idpy/LBM/MultiPhase.py:257-349; K_Collision_ShanChenMultiPhase,
    MultiPhaseKernels.py:132-216): BGK towards f_eq(n, u + F/(n omega)),
    pull streaming. Returns (f_new, n, psi, u_written) with u_written the
    mean of the pre- and post-collision velocities."""
    omega = 1.0 / tau
    xis = XI[xi]
    D = len(xis[0])
    n = f.sum(axis=0)
    psi = psi_fn(n, psi_kind)
    F = sc_force(psi, estencil, G)
    u0 = [sum(xis[q][d] * f[q] for q in range(len(xis))) / n for d in range(D)]
    ueq = [u0[d] + F[d] / (n * omega) for d in range(D)]
    feq, _ = feq_guo(xi, n, ueq, [np.zeros_like(n)] * D)
    fs = (1 - omega) * f + omega * feq
    upost = [sum(xis[q][d] * fs[q] for q in range(len(xis))) / n for d in range(D)]
    fn = np.empty_like(f)
    for q, v in enumerate(xis):
        fn[q] = _shift(fs[q], tuple(-c for c in v))
    return fn, n, psi, [0.5 * (u0[d] + upost[d]) for d in range(D)]


def init_pop(xi, n, u=None):
    D = len(XI[xi][0])
    if u is None:
        u = [np.zeros_like(n)] * D
    feq, _ = feq_guo(xi, n, u, [np.zeros_like(n)] * D)
    return feq
