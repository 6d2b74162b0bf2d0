"""Reader of CHEMKIN mechanism and NASA-7 thermo text for the oracle (test infrastructure).

Restates the reference's parsing rules so that the oracle gets the same numbers the reference would:
Species::Species / read_therm (ChemData.cpp:10-114, 260-440), Reaction::Reaction, read_reaction_line and
get_auxi_info (ChemData.cpp:868-1177), atomic weights (Element.cpp).
"""
import numpy as np

R_C = 1.98720425864083       # cal/(mol K), Constants.h:19
ATOMIC_WEIGHT = {"H": 1.008, "HE": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999, "AR": 39.95}


def _tokens_until(lines, pos, stop):
    out = []
    while pos < len(lines):
        for tok in lines[pos].split():
            if tok.startswith("!"):
                break
            t = tok.upper()
            if t in stop:
                return out, pos, t
            out.append(t)
        pos += 1
    return out, pos, None


def parse(mech_text, therm_text):
    lines = mech_text.splitlines()
    pos = 0
    while pos < len(lines) and lines[pos].strip().upper().split()[:1] not in (["ELEMENTS"], ["ELEM"]):
        pos += 1
    pos += 1
    elems = []
    while True:
        toks, pos, stop = _tokens_until(lines, pos, {"END", "SPECIES", "SPEC"})
        elems += toks
        if stop in ("SPECIES", "SPEC"):
            pos += 1
            break
        pos += 1
    species = []
    while True:
        toks, pos, stop = _tokens_until(lines, pos, {"END", "REACTIONS", "REAC"})
        species += toks
        if stop in ("REACTIONS", "REAC"):
            break
        pos += 1
    ns, ne = len(species), len(elems)
    sidx = {s: i for i, s in enumerate(species)}
    mech = dict(species=species, elements=elems, ns=ns, ne=ne)
    # ---- thermo (NASA-7, two ranges; fixed columns)
    tl = therm_text.splitlines()
    p = 0
    while not tl[p].upper().startswith("THERMO"):
        p += 1
    p += 1
    while tl[p].startswith("!") or not tl[p].strip():
        p += 1
    Tl, Tm, Th = (float(v) for v in tl[p].split()[:3])
    p += 1
    t_low, t_mid, t_high = np.full(ns, Tl), np.full(ns, Tm), np.full(ns, Th)
    lo, hi = np.zeros((ns, 7)), np.zeros((ns, 7))
    comp = np.zeros((ns, ne), dtype=int)
    mw = np.zeros(ns)
    seen = set()
    while p < len(tl):
        line = tl[p]
        if line.startswith("!") or not line.strip():
            p += 1
            continue
        key = line.split()[0].upper()
        if key == "END":
            p += 1
            continue
        name = line[:18].split()[0].upper()
        if name not in sidx or name in seen:
            p += 4
            continue
        k = sidx[name]
        seen.add(name)
        t_low[k] = float(line[45:55])
        t_high[k] = float(line[55:65])
        tm = line[65:75].split()
        if tm:
            t_mid[k] = float(tm[0])
        for e in range(4):
            cs = line[24 + 5 * e:29 + 5 * e].strip()
            if not cs or cs.startswith("0"):
                break
            el = cs[:2].strip().upper()
            n = int(cs[2:].strip() or 0)
            if el in elems:
                comp[k, elems.index(el)] = n
        mw[k] = sum(ATOMIC_WEIGHT[el] * comp[k, elems.index(el)] for el in elems)
        c = []
        for q in range(1, 4):
            l2 = tl[p + q]
            for f in range(5):
                s = l2[15 * f:15 * f + 15]
                if s.strip():
                    c.append(float(s))
        hi[k] = c[0:7]
        lo[k] = c[7:14]
        p += 4
        if len(seen) == ns:
            break
    mech.update(mw=mw, t_low=t_low, t_mid=t_mid, t_high=t_high, lo=lo, hi=hi, comp=comp)
    # ---- reactions
    while pos < len(lines) and not lines[pos].strip().upper().startswith(("REACTIONS", "REAC")):
        pos += 1
    pos += 1
    R = []
    dups = []

    def new_reac():
        return dict(type=1, rev=0, nf=np.zeros(ns), nb=np.zeros(ns), order=0.0, A=0.0, b=0.0, Ea=0.0, A2=0.0,
                    b2=0.0, Ea2=0.0, tb=np.ones(ns), troe=[0.0, 0.0, 0.0, 0.0], dup=False)

    def side(text, stoi):
        order = 0
        for part in text.split("+"):
            if not part:
                continue
            n = 1
            if part[0].isdigit():
                n = int(part[0])
                part = part[1:]
            stoi[sidx[part]] += n
            order += n
        return order

    while pos < len(lines):
        raw = lines[pos]
        up = raw.upper().strip()
        pos += 1
        if not up or up.startswith("!"):
            continue
        if up.startswith("END"):
            break
        body = up.split("!")[0]
        if "=" not in body:
            continue
        toks = body.split()
        A, b, Ea = float(toks[-3]), float(toks[-2]), float(toks[-1])
        eq = "".join(toks[:-3])
        r = new_reac()
        if "(+M)" in eq:
            r["type"] = 5
            eq = eq.replace("(+M)", "")
        elif "+M" in eq:
            r["type"] = 4
            eq = eq.replace("+M", "")
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            r["type"] = 0
        else:
            lhs, rhs = eq.split("=")
        r["order"] = -side(lhs, r["nf"]) + side(rhs, r["nb"])
        r["A"], r["b"], r["Ea"] = A, b, Ea / R_C
        # auxiliary lines
        while pos < len(lines):
            aux = lines[pos].upper().replace("/", " ")
            s = aux.split("!")[0].strip()
            if not s:
                pos += 1
                continue
            if "=" in s or s.startswith("END"):
                break
            pos += 1
            kw = s.split()
            if kw[0] in ("DUPLICATE", "DUP"):
                r["dup"] = True
                r["type"] = 3
            elif kw[0] == "LOW":
                r["A2"], r["b2"], r["Ea2"] = float(kw[1]), float(kw[2]), float(kw[3]) / R_C
            elif kw[0] == "REV":
                r["A2"], r["b2"], r["Ea2"] = float(kw[1]), float(kw[2]), float(kw[3]) / R_C
                r["rev"] = 1
            elif kw[0] == "TROE":
                r["troe"][:3] = [float(kw[1]), float(kw[2]), float(kw[3])]
                r["type"] = 6
                if len(kw) > 4:
                    r["troe"][3] = float(kw[4])
                    r["type"] = 7
            elif kw[0] == "PLOG":
                continue
            else:
                for q in range(0, len(kw) - 1, 2):
                    r["tb"][sidx[kw[q]]] = float(kw[q + 1])
        if r["dup"]:
            match = [d for d in dups if np.array_equal(R[d]["nf"], r["nf"]) and np.array_equal(R[d]["nb"], r["nb"])]
            if match:
                d = match[0]
                R[d]["A2"], R[d]["b2"], R[d]["Ea2"] = r["A"], r["b"], r["Ea"]
                dups.remove(d)
                continue
            dups.append(len(R))
        R.append(r)
    nr = len(R)
    mech["nr"] = nr
    for key in ("type", "rev"):
        mech[key] = np.array([r[key] for r in R], dtype=np.int32)
    for key in ("A", "b", "Ea", "A2", "b2", "Ea2", "order"):
        mech[key] = np.array([r[key] for r in R], dtype=float)
    mech["troe"] = np.array([r["troe"] for r in R], dtype=float)
    mech["nf"] = np.array([r["nf"] for r in R])
    mech["nb"] = np.array([r["nb"] for r in R])
    mech["tb"] = np.array([r["tb"] for r in R])
    return mech
