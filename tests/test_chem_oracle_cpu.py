"""Pins of the oracle's chemistry (CPU only).

* example/constantVolumeReactor/const_volume_reactor_output.dat: the final equilibrium state of the
  reference's ROK4E reactor run (thermo, reverse rates through Gibbs energies, Newton temperature).  The
  fixture's mechanism file (setup names chemistry/combustion_mechanism_H2.inp) is not shipped, so its
  ignition time is not reproducible: the shipped Li-2004 mechanism ignites later.  Only the equilibrium
  end state is compared.
* the reference keeps two implementations of the Li-2004 mechanism: the generic, file-driven path and
  the hard-coded routine chemical_source_hardCoded1 (FiniteRateChem.cu:19-738).  The hard-coded routine
  is restated below from its source and compared with the oracle's generic path.
* conservation: total mass and element moles are invariants of every integrator step.
"""
import os

import numpy as np
import pytest

import chem_inputs as C
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
Rc, Ru, patm = 1.98720425864083, 8314.462618, 101325.0


def load_mech():
    d = C.parse(open(os.path.join(GOLDEN, "h2_li2004.inp")).read(),
                open(os.path.join(GOLDEN, "therm_h2o2n2.dat")).read())
    return d, O.mech_struct(d)


d, m = load_mech()
mw = d["mw"]


def gibbs(t):
    g=np.zeros(9)
    for i in range(9):
        tt=min(max(t,d['t_low'][i]),d['t_high'][i])
        c=d['lo'][i] if (t<d['t_low'][i] or (t<=d['t_high'][i] and t<d['t_mid'][i])) else d['hi'][i]
        g[i]=c[0]*(1-np.log(tt))-0.5*c[1]*tt-c[2]*tt**2/6-c[3]*tt**3/12-c[4]*tt**4*0.05+c[5]/tt-c[6]
    return g
def hc(t, c):
    g=gibbs(t); tp=patm/Ru*1e-3/t; iT=1/t; iR=1/Rc*iT; lT=np.log(t)
    cc=c[0]*2.5+c[1]+c[2]+c[3]+c[4]+c[5]+c[6]+c[7]*12+c[8]
    R=[]  # (kf, reactants, products, dn)
    def add(kf, re, pr): R.append((kf, re, pr))
    add(3.55e15*np.exp(-0.41*lT-1.66e4*iR),[1,2],[3,4])
    add(5.08e4*np.exp(2.67*lT-6290*iR),[0,3],[1,4])
    add(2.16e8*np.exp(1.51*lT-3430*iR),[0,4],[7,1])
    add(2.97e6*np.exp(2.02*lT-13400*iR),[3,7],[4,4])
    add(4.58e19*np.exp(-1.40*lT-1.0438e5*iR)*cc,[0],[1,1])
    add(6.16e15*np.sqrt(iT)*cc,[3,3],[2])
    add(4.71e18*iT*cc,[1,3],[4])
    add(3.8e22*iT*iT*cc,[1,4],[7])
    kh=1.48e12*np.exp(0.6*lT); kl=6.37e20*np.exp(-1.72*lT-520*iR)
    cc2=c[0]*2+c[1]+c[2]*0.78+c[3]+c[4]+c[5]+c[6]+c[7]*11+c[8]
    pr=kl*cc2/kh; lfc=np.log10(0.8); cT=-0.4-0.67*lfc; nT=0.75-1.27*lfc; lp=np.log10(pr); tm=(lp+cT)/(nT-0.14*(lp+cT))
    add(kh*pr/(1+pr)*10**(lfc/(1+tm*tm)),[1,2],[5])
    add(1.66e13*np.exp(-820*iR),[5,1],[0,2])
    add(7.08e13*np.exp(-300*iR),[5,1],[4,4])
    add(3.25e13,[5,3],[2,4])
    add(2.89e13*np.exp(500*iR),[5,4],[7,2])
    add(4.20e14*np.exp(-11980*iR)+1.30e11*np.exp(1630*iR),[5,5],[6,2])
    kh=2.95e14*np.exp(-48400*iR); kl=1.2e17*np.exp(-45500*iR); pr=kl*cc/kh; lfc=np.log10(0.5); cT=-0.4-0.67*lfc; nT=0.75-1.27*lfc; lp=np.log10(pr); tm=(lp+cT)/(nT-0.14*(lp+cT))
    add(kh*pr/(1+pr)*10**(lfc/(1+tm*tm)),[6],[4,4])
    add(2.41e13*np.exp(-3970*iR),[6,1],[7,4])
    add(4.82e13*np.exp(-7950*iR),[6,1],[5,0])
    add(9.55e6*t*t*np.exp(-3970*iR),[6,3],[4,5])
    add(1e12+5.8e14*np.exp(-9560*iR),[6,4],[5,7])
    w=np.zeros(9)
    for kf,re,prd in R:
        dg=sum(g[p] for p in prd)-sum(g[r] for r in re); dn=len(prd)-len(re)
        kb=kf*np.exp(dg)/tp**dn
        q=kf*np.prod([c[r] for r in re])-kb*np.prod([c[p] for p in prd])
        for r in re: w[r]-=q
        for p in prd: w[p]+=q
    return w * 1e3 * mw


def reactor_initial():
    ns = d["ns"]
    X = np.zeros(ns)
    X[0], X[2] = 0.295775, 0.147887
    X[ns - 1] = 1 - X[0] - X[2]
    Wbar = np.sum(X * mw)
    Y = X * mw / Wbar
    T0, p0 = 1000.0, 202650.0
    rho = p0 * Wbar / (Ru * T0)
    h, cp = O.chem_h_cp(m, T0)
    E = np.sum(Y * h) - p0 / rho
    return rho, E, Y, T0


def test_parser_reads_li2004():
    assert d["species"] == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert d["nr"] == 19  # 21 lines, two DUPLICATE pairs merged
    assert list(d["type"]) == [1, 1, 1, 1, 4, 4, 4, 4, 6, 1, 1, 1, 1, 3, 6, 1, 1, 1, 3]
    np.testing.assert_allclose(mw, [2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])


def test_generic_kinetics_matches_hardcoded_twin():
    rng = np.random.default_rng(7)
    for _ in range(20):
        T = rng.uniform(300, 3000)
        rho = rng.uniform(0.05, 5)
        Y = rng.dirichlet(np.ones(9))
        c = rho * Y / mw * 1e-3
        w_ref = hc(T, c)
        w, wd, q1, q2 = O.chem_source(m, T, rho * Y)
        assert np.max(np.abs(w - w_ref)) <= 1e-12 * np.max(np.abs(w_ref))


def test_reactor_equilibrium_matches_fixture():
    rho, E, Y, T0 = reactor_initial()
    y, T, h, n, rej, rec = O.chem_rok4e(m, rho, E, Y * rho, T0, 1.2e-3, 1e-6, 1e-12, 1e-8)
    ref = np.loadtxt(os.path.join(GOLDEN, "const_volume_reactor_output.dat"), delimiter=",", skiprows=1)[-1]
    Yk = y / rho
    imw = np.sum(Yk / mw)
    p = rho * Ru * imw * T
    X = Yk / (mw * imw)
    assert T == pytest.approx(ref[1], rel=2e-6)
    assert p == pytest.approx(ref[2], rel=2e-6)
    np.testing.assert_allclose(X, ref[3:12], rtol=2e-5, atol=1e-9)


def test_integrator_conserves_mass_and_elements():
    rho, E, Y, T0 = reactor_initial()
    y, T, h, n, rej, rec = O.chem_rok4e(m, rho, E, Y * rho, T0, 5e-4, 1e-6, 1e-12, 1e-8)
    assert np.sum(y) == pytest.approx(rho, rel=1e-12)
    comp = d["comp"]
    el0 = (Y * rho / mw) @ comp
    el1 = (y / mw) @ comp
    np.testing.assert_allclose(el1, el0, rtol=1e-10)
    # the Newton temperature reproduces the internal energy
    hk, cp = O.chem_h_cp(m, T)
    Yk = y / rho
    assert np.sum(Yk * hk) - Ru * np.sum(Yk / mw) * T == pytest.approx(E, rel=1e-7)
