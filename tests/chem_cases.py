"""Seeded chemistry states for the batched-chemistry tests and benchmark (H2/air, Li 2004 mechanism,
species order H2 H O2 O OH HO2 H2O2 H2O N2)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MECH = os.path.join(GOLDEN, "h2_li2004.inp")
THERM = os.path.join(GOLDEN, "therm_h2o2n2.dat")
TRAN_FILE = os.path.join(GOLDEN, "tran.dat")  # transport data of the H2/O2/N2 case (example/constantVolumeReactor/input/chemistry/tran.dat)
RU, PATM = 8314.462618, 101325.0
MW = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
# stoichiometric H2/air mole fractions of the reference reactor example
X_STOICH = np.array([0.295775, 0, 0.147887, 0, 0, 0, 0, 0, 1 - 0.295775 - 0.147887])


def y_from_x(X):
    return X * MW / np.sum(X * MW)


Y_STOICH = y_from_x(X_STOICH)


def density(p, T, Y):
    """Ideal-gas density of mixtures Y (n_spec, n)."""
    return p / (RU * T * np.sum(Y / MW[:, None], axis=0))


def random_states(n, seed, t_lo=300.0, t_hi=2600.0, p_lo=0.5 * PATM, p_hi=2 * PATM):
    """Stoichiometric H2/air at a random reaction progress (fresh -> H2O/N2 products) with small radical
    pools, T uniform in [t_lo, t_hi], p uniform in [p_lo, p_hi].  Returns rho (n,), T (n,), Y (9, n)."""
    rng = np.random.default_rng(seed)
    prog = rng.uniform(0, 1, n)
    # products of complete combustion: 2 H2 + O2 -> 2 H2O
    Xp = np.array([0, 0, 0, 0, 0, 0, 0, 2 * 0.295775, 1 - 0.295775 - 0.147887])
    Xp /= Xp.sum()
    Yf, Yp = Y_STOICH, y_from_x(Xp)
    Y = Yf[:, None] * (1 - prog) + Yp[:, None] * prog
    rad = rng.uniform(0, 1, (9, n)) * np.array([0, 2e-3, 0, 2e-3, 4e-3, 5e-4, 2e-4, 0, 0])[:, None] * prog
    Y = Y + rad
    Y /= Y.sum(axis=0)
    T = rng.uniform(t_lo, t_hi, n)
    p = rng.uniform(p_lo, p_hi, n)
    return density(p, T, Y), T, Y


def hot_kernel_box(nx=128, seed=3, n_kernels=16):
    """128^3-type box: stoichiometric H2/air at 1 atm, 300 K, plus Gaussian hot kernels (peak uniform in
    [1800, 2400] K, radius uniform in [4, 10] cells, centres uniform in the periodic box).  Returns rho, T,
    Y (9, nx^3) with the i index fastest."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, nx, (n_kernels, 3))
    peaks = rng.uniform(1800, 2400, n_kernels)
    radii = rng.uniform(4, 10, n_kernels)
    ax = np.arange(nx, dtype=np.float64)
    T = np.full((nx, nx, nx), 300.0)  # [k, j, i]
    for c, pk, r in zip(centres, peaks, radii):
        d = [np.minimum(np.abs(ax - c[q]), nx - np.abs(ax - c[q])) for q in range(3)]
        r2 = d[2][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[0][None, None, :] ** 2
        T += (pk - 300.0) * np.exp(-r2 / r ** 2)
    T = np.minimum(T, peaks.max()).ravel()
    Y = np.repeat(Y_STOICH[:, None], T.size, axis=1)
    return density(PATM, T, Y), T, Y


def element_moles(Y, comp):
    """Element moles per kg: comp (n_spec, n_elem), Y (n_spec, n) -> (n_elem, n)."""
    return comp.T @ (Y / MW[:, None])
