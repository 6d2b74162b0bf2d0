"""Training pipeline of the chemistry surrogate: states from the exact integrator, labels from the exact
integrator, GPU training with the library's own kernels, and the gate set from the training box.

All integration and training runs on the GPU through the C ABI (corefl.Chemistry, corefl.Surrogate); numpy
only assembles features and scaling.  Species order is the mechanism's (H2 H O2 O OH HO2 H2O2 H2O N2 for
the Li 2004 H2/air mechanism)."""
import numpy as np

import corefl

RU = 8314.462618
PATM = 101325.0


def mixture_imw(chem_mw, Y):
    return np.sum(Y / chem_mw[:, None], axis=0)


# Transforms: mass-fraction features Y^Y_POWER; targets the signed logarithm of the rates with scales
# R0 (1/s for the nine dY/dt, K/s for dT/dt), so that rates spanning many decades train evenly.
Y_POWER = 0.1
R0 = np.array([1.0] * 9 + [1e3])


def features(mw, rho, T, Y, dt, y_power=Y_POWER):
    """Physical features (n, 12): T, ln p, Y_1^a..Y_9^a, ln dt."""
    p = rho * RU * mixture_imw(mw, Y) * T
    dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), T.shape)
    Yf = Y if y_power == 1.0 else np.where(Y > 0, np.abs(Y) ** y_power, 0.0)
    return np.column_stack([T, np.log(p), Yf.T, np.log(dt)])


def signed_log(r, r0=R0):
    return np.sign(r) * np.log1p(np.abs(r) / r0)


def signed_exp(s, r0=R0):
    return np.sign(s) * r0 * np.expm1(np.abs(s))


def ignition_states(chem, n_reactors, n_snap, seed, y_fresh, t_lo=1000.0, t_hi=2400.0, p_lo=0.5 * PATM,
                    p_hi=2 * PATM, t_first=1e-8, t_last=1e-3):
    """Constant-volume ignition of n_reactors fresh mixtures (T0, p0 uniform) recorded at n_snap geometric
    times in [t_first, t_last], plus the initial states.  Returns rho, T, Y of all snapshots."""
    rng = np.random.default_rng(seed)
    mw = chem.table("mw")
    T0 = rng.uniform(t_lo, t_hi, n_reactors)
    p0 = rng.uniform(p_lo, p_hi, n_reactors)
    Y0 = np.repeat(np.asarray(y_fresh, dtype=np.float64)[:, None], n_reactors, axis=1)
    rho = p0 / (RU * T0 * mixture_imw(mw, Y0))
    chem.set_cells(rho, T0, rho * Y0)
    times = np.concatenate([[0.0], np.geomspace(t_first, t_last, n_snap)])
    Ts, Ys = [T0.copy()], [Y0.copy()]
    for j in range(n_snap):
        chem.integrate(times[j + 1] - times[j])
        g = chem.get_cells()
        Ts.append(g["T"].copy())
        Ys.append(g["rhoY"] / rho)
    return np.tile(rho, n_snap + 1), np.concatenate(Ts), np.concatenate(Ys, axis=1)


def label(chem, rho, T, Y, dt, rtol=1e-6, atol=1e-8):
    """Exact sub-step of every state over its own dt: returns (dY (9, n), dT (n,), steps (n,))."""
    chem.set_cells(rho, T, rho * Y)
    chem.set_cell_dt(dt)
    chem.integrate(0.0, rtol, atol)
    g = chem.get_cells()
    return g["rhoY"] / rho - Y, g["T"] - T, g["steps"]


class Scaling:
    def __init__(self, in_mu, in_sig, out_mu, out_sig):
        self.in_mu, self.in_sig, self.out_mu, self.out_sig = in_mu, in_sig, out_mu, out_sig

    @classmethod
    def fit(cls, F, R):
        """Mean and standard deviation per column, the deviation floored so that a column that is constant up
        to rounding (an inert species, a fixed element ratio) is not blown up into noise."""
        mu, sig = F.mean(axis=0), F.std(axis=0)
        sig = np.maximum(sig, 1e-3 * np.maximum(np.abs(mu), 1.0))
        omu, osig = R.mean(axis=0), R.std(axis=0)
        osig = np.maximum(osig, 1e-3)
        return cls(mu, sig, omu, osig)

    def apply(self, surr):
        surr.set_scaling(self.in_mu, self.in_sig, self.out_mu, self.out_sig)


def build_dataset(chem, n_states, seed, y_fresh, dt_lo=1e-9, dt_hi=1e-7, t_min=300.0, t_max=3200.0,
                  p_lo=0.5 * PATM, p_hi=2 * PATM, n_reactors=8192, n_snap=48):
    """n_states (rho, T, Y, dt) with exact labels.  Half come from constant-volume ignition trajectories of
    n_reactors fresh mixtures (T0 in [1000, 2400] K) recorded at n_snap geometric times in [1e-7, 3e-4] s,
    a quarter are fresh mixtures at T uniform in [t_min, 2600 K], a quarter lie on the mixing line between
    fresh mixture and trajectory states at T uniform in [t_min, 2600 K]; p in [p_lo, p_hi], dt log-uniform in
    [dt_lo, dt_hi]; states hotter than t_max (burnt gas reaches about 3000 K) are dropped."""
    rng = np.random.default_rng(seed)
    mw = chem.table("mw")
    r_rho, r_T, r_Y = ignition_states(chem, n_reactors, n_snap, seed + 1, y_fresh, p_lo=p_lo, p_hi=p_hi,
                                      t_first=1e-7, t_last=3e-4)
    n_a = n_states // 2
    pick = rng.choice(r_T.size, n_a, replace=r_T.size < n_a)
    Y_a, T_a = r_Y[:, pick], r_T[pick]
    p_a = r_rho[pick] * RU * mixture_imw(mw, Y_a) * T_a
    n_b = n_states // 4
    n_c = n_states - n_a - n_b
    Y_b = np.repeat(np.asarray(y_fresh, dtype=np.float64)[:, None], n_b, axis=1)
    T_b = rng.uniform(t_min, 2600.0, n_b)
    p_b = rng.uniform(p_lo, p_hi, n_b)
    lam = rng.uniform(0, 1, n_c)
    other = r_Y[:, rng.choice(r_T.size, n_c)]
    Y_c = (1 - lam) * np.asarray(y_fresh, dtype=np.float64)[:, None] + lam * other
    T_c = rng.uniform(t_min, 2600.0, n_c)
    p_c = rng.uniform(p_lo, p_hi, n_c)
    Y = np.concatenate([Y_a, Y_b, Y_c], axis=1)
    T = np.concatenate([T_a, T_b, T_c])
    p = np.clip(np.concatenate([p_a, p_b, p_c]), p_lo, p_hi)
    rho = p / (RU * T * mixture_imw(mw, Y))
    dt = np.exp(rng.uniform(np.log(dt_lo), np.log(dt_hi), T.size))
    keep = (T >= t_min) & (T <= t_max)
    rho, T, Y, dt = rho[keep], T[keep], Y[:, keep], dt[keep]
    dY, dT, steps = label(chem, rho, T, Y, dt)
    ok = steps > 0
    return dict(rho=rho[ok], T=T[ok], Y=Y[:, ok], dt=dt[ok], dY=dY[:, ok], dT=dT[ok])


def build_mixing_dataset(chem, n_states, seed, y_a, y_b, t_a, t_b, dt_lo=1e-9, dt_hi=1e-7, t_max=3200.0,
                         p_lo=0.5 * PATM, p_hi=2 * PATM, n_reactors=8192, n_snap=48):
    """n_states (rho, T, Y, dt) with exact labels for a mixing layer of two streams (mass fractions y_a, y_b at
    temperatures t_a, t_b): half from constant-volume ignition of n_reactors mixtures z y_a + (1 - z) y_b (z
    uniform) started at the mixed temperature z t_a + (1 - z) t_b plus up to 1200 K, recorded at n_snap
    geometric times in [1e-7, 3e-4] s; half on the mixing line itself at T between the mixed temperature and
    2600 K; p in [p_lo, p_hi], dt log-uniform in [dt_lo, dt_hi]."""
    rng = np.random.default_rng(seed)
    mw = chem.table("mw")
    y_a, y_b = np.asarray(y_a, dtype=np.float64), np.asarray(y_b, dtype=np.float64)

    def mixtures(n):
        z = rng.uniform(0, 1, n)
        return z * y_a[:, None] + (1 - z) * y_b[:, None], z * t_a + (1 - z) * t_b

    Y0, Tm = mixtures(n_reactors)
    T0 = Tm + rng.uniform(0, 1200.0, n_reactors)
    p0 = rng.uniform(p_lo, p_hi, n_reactors)
    rho0 = p0 / (RU * T0 * mixture_imw(mw, Y0))
    chem.set_cells(rho0, T0, rho0 * Y0)
    times = np.concatenate([[0.0], np.geomspace(1e-7, 3e-4, n_snap)])
    Ts, Ys = [T0.copy()], [Y0.copy()]
    for j in range(n_snap):
        chem.integrate(times[j + 1] - times[j])
        g = chem.get_cells()
        Ts.append(g["T"].copy())
        Ys.append(g["rhoY"] / rho0)
    r_T, r_Y, r_rho = np.concatenate(Ts), np.concatenate(Ys, axis=1), np.tile(rho0, n_snap + 1)
    n_a = n_states // 2
    pick = rng.choice(r_T.size, n_a, replace=r_T.size < n_a)
    Y_a, T_a = r_Y[:, pick], r_T[pick]
    p_a = r_rho[pick] * RU * mixture_imw(mw, Y_a) * T_a
    n_b = n_states - n_a
    Y_b, Tm_b = mixtures(n_b)
    T_b = Tm_b + rng.uniform(0, 1, n_b) * (2600.0 - Tm_b)
    p_b = rng.uniform(p_lo, p_hi, n_b)
    Y = np.concatenate([Y_a, Y_b], axis=1)
    T = np.concatenate([T_a, T_b])
    p = np.clip(np.concatenate([p_a, p_b]), p_lo, p_hi)
    rho = p / (RU * T * mixture_imw(mw, Y))
    dt = np.exp(rng.uniform(np.log(dt_lo), np.log(dt_hi), T.size))
    keep = T <= t_max
    rho, T, Y, dt = rho[keep], T[keep], Y[:, keep], dt[keep]
    dY, dT, steps = label(chem, rho, T, Y, dt)
    ok = steps > 0
    return dict(rho=rho[ok], T=T[ok], Y=Y[:, ok], dt=dt[ok], dY=dY[:, ok], dT=dT[ok])


def targets(data):
    """Rates (n, 10): dY_k/dt and dT/dt."""
    return np.column_stack([(data["dY"] / data["dt"]).T, data["dT"] / data["dt"]])


def train(surr, data, mw, steps=6000, batch=8192, lr=2e-3, lr_min=2e-5, seed=11, chunk=500, init_seed=3):
    """Fit the network on `data` (build_dataset) with a cosine learning-rate schedule; returns the scaling
    and the per-step losses."""
    F = features(mw, data["rho"], data["T"], data["Y"], data["dt"])
    R = signed_log(targets(data))
    sc = Scaling.fit(F, R)
    surr.set_params(init_params(init_seed))
    surr.set_data((F - sc.in_mu) / sc.in_sig, (R - sc.out_mu) / sc.out_sig)
    losses = []
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        rate = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * (done + 0.5 * n) / steps))
        losses.append(surr.train(batch, n, rate, seed=seed + done))
        done += n
    sc.apply(surr)
    surr.set_transform(Y_POWER, R0)
    return sc, np.concatenate(losses) if losses else np.empty(0)


def set_gate(surr, dT_max, dT_step=1e30, t_min=290.0, t_max=3200.0, p_lo=0.5 * PATM, p_hi=2 * PATM, dt_lo=1e-9,
             dt_hi=1e-7):
    """Gate = the training box, the temperature-consistency bound dT_max (K) and the largest temperature
    change per sub-step dT_step (K) left to the network."""
    surr.set_gate(t_min, t_max, p_lo, p_hi, dt_lo, dt_hi, dT_max, dT_step)


def init_params(seed, sizes=corefl.SURROGATE_SIZES):
    """He-type initialisation: W_l ~ N(0, 2/K_l) (1/K_l for the output layer), zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(len(sizes) - 1):
        k, n = sizes[l], sizes[l + 1]
        scale = np.sqrt((2.0 if l < len(sizes) - 2 else 1.0) / k)
        out.append((rng.standard_normal((k, n)) * scale, np.zeros(n)))
    return out


def errors_vs_exact(chem, rho, T, Y, dt, surr, rtol=1e-6, atol=1e-8):
    """Surrogate step and exact step of the same cells: returns (max |dY|, max |dT|, fallback count, the
    surrogate state, the exact state)."""
    chem.set_cells(rho, T, rho * Y)
    nf = surr.chem_step(chem, dt, rtol, atol)
    s = chem.get_cells()
    chem.set_cells(rho, T, rho * Y)
    chem.integrate(dt, rtol, atol)
    x = chem.get_cells()
    dy = np.max(np.abs(s["rhoY"] - x["rhoY"]) / rho) if T.size else 0.0
    dT = np.max(np.abs(s["T"] - x["T"])) if T.size else 0.0
    return dy, dT, nf, s, x
