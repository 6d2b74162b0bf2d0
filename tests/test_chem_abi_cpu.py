"""The library's mechanism parser (host side of the chemistry C ABI, no GPU) against the test parser of
the same CHEMKIN / NASA-7 files; both restate ChemData.cpp (Species and Reaction readers)."""
import os

import numpy as np
import pytest

import chem_inputs as C
import corefl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MECH = os.path.join(GOLDEN, "h2_li2004.inp")
THERM = os.path.join(GOLDEN, "therm_h2o2n2.dat")


@pytest.fixture(scope="module")
def both():
    d = C.parse(open(MECH).read(), open(THERM).read())
    return d, corefl.Chemistry(MECH, THERM)


def test_counts_and_names(both):
    d, ch = both
    assert (ch.n_spec, ch.n_reac) == (9, 19)
    assert ch.species == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert d["nr"] == 19


@pytest.mark.parametrize("name", ["mw", "t_low", "t_mid", "t_high", "lo", "hi", "A", "b", "Ea", "A2", "b2", "Ea2",
                                  "order", "type", "rev", "troe", "nf", "nb", "tb"])
def test_tables_match(both, name):
    d, ch = both
    np.testing.assert_array_equal(ch.table(name), np.asarray(d[name], dtype=np.float64).reshape(ch.table(name).shape))


def test_element_composition(both):
    _, ch = both
    comp = ch.table("comp")
    el = ch.elements
    h, o, n = el.index("H"), el.index("O"), el.index("N")
    assert comp[ch.species.index("H2O2"), h] == 2 and comp[ch.species.index("H2O2"), o] == 2
    assert comp[ch.species.index("N2"), n] == 2 and comp[ch.species.index("HO2"), o] == 2


def test_in_memory_create_and_errors():
    ch = corefl.Chemistry(mech_text=open(MECH).read(), therm_text=open(THERM).read())
    assert ch.n_reac == 19
    with pytest.raises(corefl.CoreflError):
        corefl.Chemistry(mech_text="ELEMENTS H END\nSPECIES H2 XX END\n", therm_text="")
    with pytest.raises(corefl.CoreflError):
        corefl.Chemistry("/nonexistent.inp", None)
