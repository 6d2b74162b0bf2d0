// CHEMKIN mechanism and NASA-7 thermo parser.  Rules follow the reference reader: element and species
// lists (ChemData.cpp:10-114), two-range NASA-7 records at fixed columns with the molecular weight from
// the element composition (ChemData.cpp:260-440, Element.cpp), reactions with +M / (+M) / LOW / TROE /
// REV / DUPLICATE and third-body efficiencies, Ea converted from cal/mol to K, DUPLICATE pairs merged into
// one reaction of type 3 (ChemData.cpp:868-1177).
#include "chem.hpp"

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace corefl {

namespace {

constexpr double kRc = 1.98720425864083; // cal/(mol K), Constants.h:19

double atomic_weight(const std::string &el) {
  static const std::map<std::string, double> w{{"H", 1.008},   {"HE", 4.002602}, {"C", 12.011},
                                                {"N", 14.007}, {"O", 15.999},    {"AR", 39.95}};
  auto it = w.find(el);
  if (it == w.end()) throw std::runtime_error("unknown element " + el);
  return it->second;
}

std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

std::string strip_comment(const std::string &s) {
  const auto p = s.find('!');
  return p == std::string::npos ? s : s.substr(0, p);
}

std::vector<std::string> lines_of(const std::string &text) {
  std::vector<std::string> out;
  std::istringstream in(text);
  std::string l;
  while (std::getline(in, l)) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    out.push_back(l);
  }
  return out;
}

struct Side {
  std::vector<std::pair<int, int>> sp; // (species, coefficient)
  int total = 0;
};

Side parse_side(const std::string &text, const std::map<std::string, int> &sidx) {
  Side s;
  std::string part;
  auto flush = [&]() {
    if (part.empty()) return;
    int n = 1;
    std::string name = part;
    if (std::isdigit((unsigned char)name[0])) {
      n = name[0] - '0';
      name = name.substr(1);
    }
    auto it = sidx.find(name);
    if (it == sidx.end()) throw std::runtime_error("reaction names unknown species " + name);
    bool merged = false;
    for (auto &p : s.sp)
      if (p.first == it->second) {
        p.second += n;
        merged = true;
      }
    if (!merged) s.sp.emplace_back(it->second, n);
    s.total += n;
    part.clear();
  };
  for (char c : text) {
    if (c == '+') flush();
    else part += c;
  }
  flush();
  return s;
}

} // namespace

ChemMech parse_mechanism(const std::string &mech_text, const std::string &therm_text) {
  ChemMech m;
  std::memset(&m, 0, sizeof(m));
  const auto L = lines_of(mech_text);
  size_t pos = 0;
  // ---- ELEMENTS / SPECIES
  std::vector<std::string> elems, specs;
  int section = 0; // 0 before, 1 elements, 2 species, 3 reactions
  for (; pos < L.size() && section < 3; ++pos) {
    std::istringstream ls(strip_comment(L[pos]));
    std::string tok;
    while (ls >> tok) {
      const std::string t = upper(tok);
      if (t == "ELEMENTS" || t == "ELEM") section = 1;
      else if (t == "SPECIES" || t == "SPEC") section = 2;
      else if (t == "REACTIONS" || t == "REAC") { section = 3; break; }
      else if (t == "END") continue;
      else if (section == 1) elems.push_back(t);
      else if (section == 2) specs.push_back(t);
    }
  }
  if (section != 3) throw std::runtime_error("mechanism has no REACTIONS section");
  if ((int)specs.size() > kMaxSpec || (int)elems.size() > kMaxElem)
    throw std::runtime_error("too many species or elements for the chemistry tables");
  m.ns = (int)specs.size();
  m.ne = (int)elems.size();
  std::map<std::string, int> sidx;
  for (int i = 0; i < m.ns; ++i) {
    sidx[specs[i]] = i;
    std::strncpy(m.species[i], specs[i].c_str(), 15);
  }
  for (int e = 0; e < m.ne; ++e) std::strncpy(m.elements[e], elems[e].c_str(), 3);

  // ---- thermo
  const auto T = lines_of(therm_text);
  size_t p = 0;
  while (p < T.size() && upper(T[p]).rfind("THERMO", 0) != 0) ++p;
  if (p == T.size()) throw std::runtime_error("thermo text has no THERMO header");
  ++p;
  while (p < T.size() && (T[p].empty() || T[p][0] == '!')) ++p;
  double Tl = 300, Tm = 1000, Th = 5000;
  {
    std::istringstream ls(T[p]);
    ls >> Tl >> Tm >> Th;
    ++p;
  }
  for (int i = 0; i < m.ns; ++i) {
    m.t_low[i] = Tl;
    m.t_mid[i] = Tm;
    m.t_high[i] = Th;
  }
  std::vector<bool> seen(m.ns, false);
  int n_seen = 0;
  while (p + 3 < T.size() && n_seen < m.ns) {
    const std::string &l = T[p];
    if (l.empty() || l[0] == '!' || l.find_first_not_of(" \t") == std::string::npos) { ++p; continue; }
    std::istringstream ls(l.substr(0, 18));
    std::string name;
    ls >> name;
    name = upper(name);
    if (name == "END" || !sidx.count(name) || seen[sidx[name]] || l.size() < 65) { p += name == "END" ? 1 : 4; continue; }
    const int k = sidx[name];
    seen[k] = true;
    ++n_seen;
    m.t_low[k] = std::stod(l.substr(45, 10));
    m.t_high[k] = std::stod(l.substr(55, 10));
    if (l.size() > 65) {
      std::istringstream tm(l.substr(65, 10));
      double v;
      if (tm >> v) m.t_mid[k] = v;
    }
    double w = 0;
    for (int e = 0; e < 4; ++e) {
      if (l.size() < (size_t)(29 + 5 * e)) break;
      std::string cs = l.substr(24 + 5 * e, 5);
      cs.erase(0, cs.find_first_not_of(' '));
      if (cs.empty() || cs[0] == '0') break;
      std::string el = upper(cs.substr(0, 2));
      el.erase(el.find_last_not_of(' ') + 1);
      const int n = std::atoi(cs.substr(2).c_str());
      auto it = std::find(elems.begin(), elems.end(), el);
      if (it == elems.end()) continue;
      m.comp[k][it - elems.begin()] = n;
    }
    for (int e = 0; e < m.ne; ++e) w += atomic_weight(elems[e]) * m.comp[k][e];
    m.mw[k] = w;
    std::vector<double> c;
    for (int q = 1; q <= 3; ++q)
      for (int f = 0; f < 5; ++f) {
        const std::string &lq = T[p + q];
        if (lq.size() < (size_t)(15 * f + 15)) continue;
        const std::string s = lq.substr(15 * f, 15);
        if (s.find_first_not_of(' ') != std::string::npos) c.push_back(std::stod(s));
      }
    if (c.size() < 14) throw std::runtime_error("short NASA-7 record for " + name);
    for (int j = 0; j < 7; ++j) {
      m.hi[k][j] = c[j];
      m.lo[k][j] = c[7 + j];
    }
    p += 4;
  }
  if (n_seen < m.ns) throw std::runtime_error("thermo data missing for some species");

  // ---- reactions
  std::vector<int> open_dups;
  int nr = 0;
  while (pos < L.size()) {
    const std::string raw = upper(strip_comment(L[pos]));
    ++pos;
    std::istringstream ls(raw);
    std::vector<std::string> tok;
    std::string t;
    while (ls >> t) tok.push_back(t);
    if (tok.empty()) continue;
    if (tok[0] == "END") break;
    if (raw.find('=') == std::string::npos) continue;
    if (tok.size() < 4) throw std::runtime_error("reaction line needs A, b, Ea: " + raw);
    if (nr >= kMaxReac) throw std::runtime_error("too many reactions for the chemistry tables");
    const int r = nr;
    std::string eq;
    for (size_t q = 0; q + 3 < tok.size(); ++q) eq += tok[q];
    m.A[r] = std::stod(tok[tok.size() - 3]);
    m.b[r] = std::stod(tok[tok.size() - 2]);
    m.Ea[r] = std::stod(tok[tok.size() - 1]) / kRc;
    m.type[r] = 1;
    m.rev[r] = 0;
    for (int s = 0; s < m.ns; ++s) m.tb[r][s] = 1.0;
    for (int q = 0; q < 4; ++q) m.troe[r][q] = 0;
    m.A2[r] = m.b2[r] = m.Ea2[r] = 0;
    auto erase_all = [&](const std::string &what) {
      for (size_t at; (at = eq.find(what)) != std::string::npos;) eq.erase(at, what.size());
    };
    if (eq.find("(+M)") != std::string::npos) {
      m.type[r] = 5;
      erase_all("(+M)");
    } else if (eq.find("+M") != std::string::npos) {
      m.type[r] = 4;
      erase_all("+M");
    }
    std::string lhs, rhs;
    if (auto a = eq.find("<=>"); a != std::string::npos) {
      lhs = eq.substr(0, a);
      rhs = eq.substr(a + 3);
    } else if (auto b2 = eq.find("=>"); b2 != std::string::npos) {
      lhs = eq.substr(0, b2);
      rhs = eq.substr(b2 + 2);
      m.type[r] = 0;
    } else {
      const auto e = eq.find('=');
      lhs = eq.substr(0, e);
      rhs = eq.substr(e + 1);
    }
    const Side f = parse_side(lhs, sidx), bk = parse_side(rhs, sidx);
    if ((int)f.sp.size() > kMaxSide || (int)bk.sp.size() > kMaxSide) throw std::runtime_error("too many species in a reaction");
    m.nfs[r] = (int)f.sp.size();
    m.nbs[r] = (int)bk.sp.size();
    for (int s = 0; s < kMaxSpec; ++s) m.nf[r][s] = m.nb[r][s] = 0;
    for (int q = 0; q < m.nfs[r]; ++q) {
      m.fsp[r][q] = f.sp[q].first;
      m.fnu[r][q] = f.sp[q].second;
      m.nf[r][f.sp[q].first] = f.sp[q].second;
    }
    for (int q = 0; q < m.nbs[r]; ++q) {
      m.bsp[r][q] = bk.sp[q].first;
      m.bnu[r][q] = bk.sp[q].second;
      m.nb[r][bk.sp[q].first] = bk.sp[q].second;
    }
    m.order[r] = bk.total - f.total;
    bool dup = false;
    // auxiliary lines
    while (pos < L.size()) {
      std::string aux = upper(strip_comment(L[pos]));
      std::replace(aux.begin(), aux.end(), '/', ' ');
      std::istringstream as(aux);
      std::vector<std::string> kw;
      while (as >> t) kw.push_back(t);
      if (kw.empty()) { ++pos; continue; }
      if (aux.find('=') != std::string::npos || kw[0] == "END") break;
      ++pos;
      if (kw[0] == "DUPLICATE" || kw[0] == "DUP") {
        dup = true;
        m.type[r] = 3;
      } else if (kw[0] == "LOW" || kw[0] == "REV") {
        if (kw.size() < 4) throw std::runtime_error(kw[0] + " needs three values");
        m.A2[r] = std::stod(kw[1]);
        m.b2[r] = std::stod(kw[2]);
        m.Ea2[r] = std::stod(kw[3]) / kRc;
        if (kw[0] == "REV") m.rev[r] = 1;
      } else if (kw[0] == "TROE") {
        for (int q = 0; q < 3 && q + 1 < (int)kw.size(); ++q) m.troe[r][q] = std::stod(kw[q + 1]);
        m.type[r] = 6;
        if (kw.size() > 4) {
          m.troe[r][3] = std::stod(kw[4]);
          m.type[r] = 7;
        }
      } else if (kw[0] == "PLOG") {
        continue;
      } else {
        for (size_t q = 0; q + 1 < kw.size(); q += 2) {
          auto it = sidx.find(kw[q]);
          if (it == sidx.end()) throw std::runtime_error("third-body efficiency for unknown species " + kw[q]);
          m.tb[r][it->second] = std::stod(kw[q + 1]);
        }
      }
    }
    if (dup) {
      auto same = [&](int a) {
        if (m.nfs[a] != m.nfs[r] || m.nbs[a] != m.nbs[r]) return false;
        for (int q = 0; q < m.nfs[r]; ++q)
          if (m.fsp[a][q] != m.fsp[r][q] || m.fnu[a][q] != m.fnu[r][q]) return false;
        for (int q = 0; q < m.nbs[r]; ++q)
          if (m.bsp[a][q] != m.bsp[r][q] || m.bnu[a][q] != m.bnu[r][q]) return false;
        return true;
      };
      auto it = std::find_if(open_dups.begin(), open_dups.end(), same);
      if (it != open_dups.end()) {
        m.A2[*it] = m.A[r];
        m.b2[*it] = m.b[r];
        m.Ea2[*it] = m.Ea[r];
        open_dups.erase(it);
        continue; // merged into the first of the pair
      }
      open_dups.push_back(r);
    }
    ++nr;
  }
  m.nr = nr;
  for (int k = 0; k < m.ns; ++k) {
    m.imw[k] = 1.0 / m.mw[k];
    m.ln_t_low[k] = std::log(m.t_low[k]);
    m.ln_t_high[k] = std::log(m.t_high[k]);
    m.inv_t_low[k] = 1.0 / m.t_low[k];
    m.inv_t_high[k] = 1.0 / m.t_high[k];
  }
  m.need_gibbs = 0;
  for (int r = 0; r < nr; ++r) m.need_gibbs |= m.type[r] != 0 && m.rev[r] != 1;
  for (int r = 0; r < nr; ++r) {
    m.umask[r] = m.amask[r] = 0;
    for (int s = 0; s < m.ns; ++s) {
      m.dnu[r][s] = m.nb[r][s] - m.nf[r][s];
      if (m.nf[r][s] != 0 || m.nb[r][s] != 0) m.umask[r] |= 1u << s;
      if (m.dnu[r][s] != 0) m.amask[r] |= 1u << s;
    }
    m.iorder[r] = (int)m.order[r];
    if ((double)m.iorder[r] != m.order[r]) throw std::runtime_error("non-integer reaction order");
  }
  return m;
}

void parse_transport(const std::string &tran_text, ChemMech &m) {
  const int ns = m.ns;
  std::vector<double> eps(ns, 0), sig(ns, 0), dip(ns, 0), pol(ns, 0);
  std::vector<bool> seen(ns, false);
  int n_read = 0;
  std::istringstream in(tran_text);
  std::string line;
  while (std::getline(in, line)) {
    for (auto &c : line) c = (char)std::toupper((unsigned char)c);
    if (line.empty() || line[0] == '!') continue;
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key)) continue;
    if (key.rfind("END", 0) == 0 || key.rfind("REAC", 0) == 0) break;
    int k = -1;
    for (int q = 0; q < ns; ++q)
      if (key == m.species[q]) k = q;
    if (k < 0 || seen[k]) continue; // unknown species, or a second record of a species read already
    double z = 0;
    if (!(ls >> m.geometry[k] >> eps[k] >> sig[k] >> dip[k] >> pol[k] >> z))
      throw std::runtime_error("transport file: bad record for " + key);
    m.lj_inv[k] = 1.0 / eps[k];
    m.vis_coeff[k] = 2.6693e-6 * std::sqrt(m.mw[k]) / (sig[k] * sig[k]);
    const double pi = 3.14159265358979323846, tri = eps[k] / 298;
    const double f298 = 1 + 0.5 * pi * pi * pi * std::sqrt(tri) + (2 + 0.25 * pi * pi) * tri +
                        std::sqrt(pi * pi * pi) * std::sqrt(tri * tri * tri);
    m.zrot[k] = z * f298;
    seen[k] = true;
    ++n_read;
  }
  if (n_read < ns) throw std::runtime_error("transport file: not every species has transport data");
  // reduced dipole moment (Debye, Angstrom, K) and the polar / non-polar correction of the pair potential
  const double kb = 1.380649e-23;
  auto mu_star = [&](int q) {
    if (dip[q] == 0) return 0.0;
    const double conv = 1e-18 * std::sqrt(1e-7) * std::pow(1e+8, 1.5) / std::sqrt(kb);
    return conv * dip[q] / std::sqrt(eps[q] * sig[q] * sig[q] * sig[q]);
  };
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      m.wdw[i][j] = std::pow(m.mw[j] / m.mw[i], 0.25);
      m.sqw[i][j] = 1.0 / std::sqrt(8 * (1 + m.mw[i] / m.mw[j]));
    }
  for (int j = 0; j < ns; ++j) {
    m.bdc[j][j] = 0.0188324 / (std::sqrt(m.mw[j] * 0.5) * sig[j] * sig[j]);
    m.kbe[j][j] = 1.0 / eps[j];
    for (int k = j + 1; k < ns; ++k) {
      const double w = m.mw[j] * m.mw[k] / (m.mw[j] + m.mw[k]);
      double xi = 1;
      if ((dip[j] != 0) != (dip[k] != 0)) {
        const int p = dip[j] != 0 ? j : k, n = dip[j] != 0 ? k : j;
        const double an = pol[n] / std::pow(sig[n], 3), mp = mu_star(p);
        xi = 1.0 + 0.25 * an * mp * mp * std::sqrt(eps[p] / eps[n]);
      }
      const double sjk = 0.5 * (sig[j] + sig[k]) * std::pow(xi, -1.0 / 6.0);
      m.bdc[j][k] = m.bdc[k][j] = 0.0188324 / (std::sqrt(w) * sjk * sjk);
      m.kbe[j][k] = m.kbe[k][j] = 1.0 / (xi * xi * std::sqrt(eps[j] * eps[k]));
    }
  }
  for (int j = 0; j < ns; ++j) {
    m.log_lj_inv[j] = std::log(m.lj_inv[j]);
    for (int k = 0; k < ns; ++k) {
      m.log_kbe[j][k] = std::log(m.kbe[j][k]);
      m.inv_bdc[j][k] = 1.0 / m.bdc[j][k];
    }
  }
  m.has_transport = 1;
}

} // namespace corefl
