"""Seeded synthetic initial densities of the benchmark workloads.

Each generator returns a float64 density array in lexicographic order
(x fastest), to be set with SetDensityField (u = 0, f = f_eq(n, 0)). Seeds:
numpy.random.default_rng(20261202 + workload number).

  droplet2d_noise  D2Q9 128^2, tanh droplet R=20 width 2 between n_g=0.1 and
                   n_l=2.0, centred at L//2, plus uniform noise +-1e-4
  droplets2d       D2Q9 4096^2, 64 droplets with radii uniform in [10, 60]
                   at random non-overlapping periodic positions
  sphere3d         D3Q19 256^3, one sphere R=40 at the centre, width 2
"""
import numpy as np

N_G, N_L = 0.1, 2.0
SEED0 = 20261202

# Coupling of the benchmark workloads. Their "G = -5" is the textbook
# Shan-Chen value in the convention where the E4 weights sum to c2 = 1/3
# (critical coupling -4 for psi = 1 - exp(-n)). idpy's FStencils are
# normalised to e2 = sum_e w_e e_x^2 = 1, which puts the critical coupling at
# -4 c2 = -4/3; the same physics is SC_G = -5 c2 = -5/3 there (liquid/gas
# densities ~1.91/0.125 for the 128^2 droplet). SC_G = -5 with e2 = 1 is far
# past the spinodal and diverges within a few steps on any implementation.
SC_G = -5.0 / 3.0


def _grid(dims):
    axes = np.meshgrid(*[np.arange(L, dtype=np.float64) for L in dims], indexing='ij')
    return [a.transpose(tuple(range(len(dims)))[::-1]) for a in axes]


def tanh_profile(r, R, W, n_g=N_G, n_l=N_L):
    return 0.5 * (n_g + n_l) - 0.5 * (n_l - n_g) * np.tanh((r - R) / W)


def droplet(dims, R, W=2.0, n_g=N_G, n_l=N_L, centre=None):
    xs = _grid(dims)
    c = [L // 2 for L in dims] if centre is None else centre
    r = np.sqrt(sum((x - cc) ** 2 for x, cc in zip(xs, c)))
    return tanh_profile(r, R, W, n_g, n_l).ravel()


def droplet2d_noise(L=128, R=20.0, W=2.0, noise=1e-4, seed=SEED0 + 1):
    rng = np.random.default_rng(seed)
    n = droplet((L, L), R, W)
    return n + rng.uniform(-noise, noise, size=n.shape)


def _periodic_dist(a, b, L):
    d = np.abs(np.asarray(a) - np.asarray(b))
    d = np.minimum(d, L - d)
    return np.sqrt((d ** 2).sum())


def droplets2d(L=4096, count=64, rmin=10.0, rmax=60.0, W=2.0, gap=4.0, seed=SEED0 + 2):
    """64 droplets: radii U[rmin, rmax], centres rejection-sampled so that
    periodic distances exceed R_i + R_j + gap; n = n_g + sum_i (n_l - n_g)
    (1 - tanh((r_i - R_i)/W)) / 2, clipped to [n_g, n_l]."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, size=count)
    centres = []
    for R in radii:
        for _ in range(100000):
            c = rng.uniform(0, L, size=2)
            if all(_periodic_dist(c, c2, L) > R + R2 + gap for c2, R2 in centres):
                centres.append((c, R))
                break
        else:
            raise RuntimeError('could not place droplets')
    n = np.full((L, L), N_G)
    x = np.arange(L, dtype=np.float64)
    for c, R in centres:
        # only a window of 2*(R+8) around each centre contributes
        h = int(R + 8 * W) + 1
        ix = (np.arange(int(c[0]) - h, int(c[0]) + h + 1)) % L
        iy = (np.arange(int(c[1]) - h, int(c[1]) + h + 1)) % L
        dx = x[ix] - c[0]
        dx = dx - L * np.round(dx / L)
        dy = x[iy] - c[1]
        dy = dy - L * np.round(dy / L)
        r = np.sqrt(dy[:, None] ** 2 + dx[None, :] ** 2)
        n[np.ix_(iy, ix)] += (N_L - N_G) * 0.5 * (1.0 - np.tanh((r - R) / W))
    return np.clip(n, N_G, N_L).ravel()


def sphere3d(L=256, R=40.0, W=2.0):
    return droplet((L, L, L), R, W)
