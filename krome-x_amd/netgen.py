"""Network kernel generator: a `.knet` network -> one C++/HIP header.

The header holds, for one network and one thermal configuration, the
functions the per-cell integrator evaluates, each a template on the real
type R (double for the FP64 path, float for the FP32 Jacobian of the
mixed-precision path) and callable from host code (the CPU oracle, g++) and
from gfx950 kernels (hipcc):

  rates_T(Tgas, p, k)         rates of the rate temperature and parameters
                              (KROME-X `rates_fixed`, codegen.py:4760-4793)
  rates_n(Tgas, n, p, k)      density-dependent rates (`rates_var`)
  rhs(n, k, dn)               mass-action dn/dt in reaction order
                              (codegen.py:4289-4327)
  jac_species(n, k, J)        analytic dn/dt Jacobian at fixed rates
                              (mass action; codegen.py:5180-5225 without
                              the dk/dn aggregate term, which the networks
                              generated here with an analytic Jacobian do
                              not have: every rate is a function of T)
  thermal(Tth, n, k, p)       dTgas/dt = (gamma-1)(heat-cool)/(kB ntot)
                              (solve_cell.hpp:305-429, the configured
                              terms only)

The semantics follow the KROME-X network format (docs/network-format.md):
species in code-point order (section 2.1), half-open T pieces tiled into
piecewise rates with the `asym` outside rule (zero below, clamped above;
section 3.1), the rate language of section 4 (`var`, `extern`, built-in
functions, `n(X)`, `Tgas_eV`, `ntot`, `Hnuclei`).  The parser and the
emitter are this project's own; only the format, the element masses
(kromex/knet.py:64-77) and the H2 chemistry energy ledger
(codegen.py:4953-4961) are data that must agree with the reference.
"""
from __future__ import annotations

import ast
import math
import re
from dataclasses import dataclass, field
from pathlib import Path

# element masses [amu] and the electron mass: the reference's table
# (kromex/knet.py:64-77), so species masses (rho, tff, x <-> n) agree
MASS_AMU = {
    "H": 1.008, "D": 2.014, "He": 4.0026, "Li": 6.94, "Be": 9.012,
    "B": 10.81, "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998,
    "Ne": 20.180, "Na": 22.990, "Mg": 24.305, "Al": 26.982, "Si": 28.085,
    "P": 30.974, "S": 32.06, "Cl": 35.45, "Ar": 39.948, "K": 39.098,
    "Ca": 40.078, "Sc": 44.956, "Ti": 47.867, "V": 50.942, "Cr": 51.996,
    "Mn": 54.938, "Fe": 55.845, "Co": 58.933, "Ni": 58.693, "Cu": 63.546,
    "Zn": 65.38,
}
ELECTRON_MASS_AMU = 5.48579909e-4
AMU_G = 1.6605390666e-24          # CODATA 2018
BOLTZMANN_EV = 8.617333262145179e-05
BOLTZMANN_ERG = 1.380649e-16
EV_TO_ERG = 1.602176634e-12
CONSTANTS = {"boltzmann_eV": BOLTZMANN_EV, "boltzmann_erg": BOLTZMANN_ERG,
             "eV_to_erg": EV_TO_ERG, "pi": math.pi,
             "N_avogadro": 6.02214076e+23, "planck_eV": 4.135667696923859e-15,
             "Rgas_kJ": 0.00831446261815324}
# default values of the predefined parameters (kromex/knet.py:53-58)
PARAM_DEFAULTS = {"gamma_ad": 5.0 / 3.0, "Tcmb": 2.725, "tff": math.nan,
                  "opratio": 3.0, "G0": 0.0, "d2g": 0.01, "Av": 0.0,
                  "zredshift": 0.0, "Tdust": 10.0}
_ELEMENTS_BY_LEN = sorted(MASS_AMU, key=len, reverse=True)
_SPECIES_RE = re.compile(r"^(?P<curly>\{)?(?P<core>[A-Za-z][A-Za-z0-9]*)"
                         r"(?P<labels>(?:_[A-Za-z0-9]+)*)(?P<charge>\++|-+)?"
                         r"(?P<curly2>\})?$")
# the extern functions this build provides (kx_physics.h): name ->
# (the densities they read: None, "agg" for an aggregate (every species)
# or one species; the parameters they read; their C++ call, with the
# knet arguments as {0}, {1} ... and the e- index as {e}).  H2_sigmaLW is
# a cross section, called only inside photo(xsec=...)
EXTERNS = {
    "dissH2_Martin96": ("agg", (), "kx::dissH2_Martin96(Tgas, Hnuclei(n))"),
    "H2_solomonLW": (None, (), "kx::H2_solomonLW({0})"),
    "H_recombination_on_dust": (
        "e-", ("d2g", "G0"),
        "kx::H_recombination_on_dust(Tgas, n[{e}], R(p.G0), R(p.d2g))"),
    "He_recombination_on_dust": (
        "e-", ("d2g", "G0"),
        "kx::He_recombination_on_dust(Tgas, n[{e}], R(p.G0), R(p.d2g))"),
    "H2_sigmaLW": (None, (), None),
    "sigma_v96": (None, (), None),
    "sigma_vy95": (None, (), None),
    # the legacy rate-database fits (codegen.py:826-852): functions of Tgas
    "colion_v96": (None, (), "kx::colion_v96({0}, {1}, {2}, {3}, {4}, {5})"),
    "recV96": (None, (), "kx::recV96({0}, {1}, {2})"),
    "recNewV96": (None, (), "kx::recNewV96({0}, {1}, {2}, {3}, {4})"),
    "recFeV96": (None, (), "kx::recFeV96({0}, {1}, {2}, {3})"),
    # H2 self-shielding at the default options (radiation.shielding WG11,
    # num2col default): a function of Tgas and n(H2) (codegen.py:1903-1939)
    "krome_fshield": ("H2", (), "kx::fshield_wg11(Tgas, n[{h2}])"),
}
# the predefined parameters in the order the parameter rows list them,
# after the network's own (sorted): kromex/codegen.py:287-312
UNIVERSAL_PARAMS = ("Tdust", "G0", "d2g", "Av", "gamma_ad", "Tcmb", "tff",
                    "zredshift", "opratio")


class KnetError(Exception):
    pass


@dataclass
class Species:
    name: str
    comp: dict
    charge: int
    mass_amu: float
    # a name that is not an element composition (GRAIN0, a generic
    # collider): no atoms, zero mass before its charge's electrons
    # (docs/network-format.md:60-66)
    opaque: bool = False

    @property
    def cid(self):
        """C identifier: + -> j, - -> k (KROME's IDX_ names)"""
        if self.name == "e-":
            return "E"
        return self.name.replace("+", "j").replace("-", "k")


def parse_species(tok):
    """a species token -> (Species, kinetic)"""
    tok = tok.strip()
    if tok == "e-":
        return Species("e-", {}, -1, ELECTRON_MASS_AMU), True
    m = _SPECIES_RE.match(tok)
    if not m or bool(m.group("curly")) != bool(m.group("curly2")):
        raise KnetError("malformed species %r" % tok)
    core = m.group("core")
    comp = _composition(core)
    ch = m.group("charge") or ""
    charge = len(ch) * (1 if ch.startswith("+") else -1) if ch else 0
    opaque = comp is None
    if opaque:
        comp = {}
    mass = sum(MASS_AMU[e] * c for e, c in comp.items())
    name = core + (m.group("labels") or "") + ch
    return (Species(name, comp, charge, mass - charge * ELECTRON_MASS_AMU,
                    opaque), not m.group("curly"))


def _composition(core):
    """{element: atoms} of a species core read greedily, the longest
    element symbol first; None when some part is not an element symbol"""
    comp, i = {}, 0
    while i < len(core):
        el = next((e for e in _ELEMENTS_BY_LEN if core.startswith(e, i)),
                  None)
        if el is None:
            return None
        i += len(el)
        j = i
        while j < len(core) and core[j].isdigit():
            j += 1
        comp[el] = comp.get(el, 0) + int(core[i:j] or 1)
        i = j
    return comp


@dataclass
class Reaction:
    reactants: list          # [(name, kinetic)]
    products: list           # [name]
    tlo: float
    thi: float
    ranged: bool
    rate: str
    line: int
    index: int = -1          # k[] slot, in file order
    group: int = -1          # piecewise group id

    @property
    def eq(self):
        lhs = " + ".join(("{%s}" % s) if not k else s
                         for s, k in self.reactants)
        return "%s -> %s" % (lhs, " + ".join(self.products))


@dataclass
class Network:
    name: str
    species: list = field(default_factory=list)   # Species, code-point order
    reactions: list = field(default_factory=list)
    vars: dict = field(default_factory=dict)      # name -> expr text
    params: dict = field(default_factory=dict)    # network params
    externs: list = field(default_factory=list)

    def idx(self, name):
        for i, s in enumerate(self.species):
            if s.name == name:
                return i
        raise KnetError("unknown species %r" % name)

    @property
    def nspec(self):
        return len(self.species)

    @property
    def nrea(self):
        return len(self.reactions)


def _number(tok):
    t = tok.strip().lower()
    if t in ("inf", "+inf", "infinity"):
        return math.inf
    return float(t)


def parse_knet(path):
    """parse a `.knet` file into a Network (docs/network-format.md 1-4)"""
    path = Path(path)
    text = path.read_text()
    # continuation lines, comments
    lines, buf, start = [], "", None
    for i, raw in enumerate(text.splitlines(), 1):
        s = raw.split("#", 1)[0].rstrip()
        if start is None:
            start = i
        if s.endswith("\\"):
            buf += s[:-1] + " "
            continue
        buf += s
        if buf.strip():
            lines.append((start, buf.strip()))
        buf, start = "", None
    if not lines or not lines[0][1].startswith("krome_format"):
        raise KnetError("%s: no krome_format header" % path)
    net = Network(name=path.stem)
    spec = {}
    rx = []
    for ln, s in lines[1:]:
        head = s.split(None, 1)[0]
        if head == "var":
            name, expr = s[3:].split("=", 1)
            net.vars[name.strip()] = expr.strip()   # later definitions win
            continue
        if head == "param":
            name, val = s[5:].split("=", 1)
            net.params[name.strip()] = float(val)
            continue
        if head == "extern":
            for e in s[6:].replace(",", " ").split():
                if e not in EXTERNS:
                    raise KnetError("line %d: extern %s is not provided"
                                    % (ln, e))
                net.externs.append(e)
            continue
        if head == "species":
            for t in s[7:].replace(",", " ").split():
                sp, _ = parse_species(t)
                spec[sp.name] = sp
            continue
        if head in ("ref",):
            continue
        if head in ("table", "cooling", "heating", "rate_scale", "freeze"):
            raise KnetError("line %d: directive %r is not supported by this "
                            "generator" % (ln, head))
        if "->" not in s:
            raise KnetError("line %d: not a reaction: %r" % (ln, s))
        lhs, rest = s.split("->", 1)
        b = rest.index("[")
        e = rest.index("]", b)
        prods = [t for t in rest[:b].split() if t != "+"]
        rng = rest[b + 1:e].strip()
        rate = rest[e + 1:].split(";", 1)[0].strip()
        if rng:
            lo, hi = (x.strip() for x in rng.split(","))
            tlo = _number(lo) if lo else 0.0
            thi = _number(hi) if hi else math.inf
            ranged = True
        else:
            tlo, thi, ranged = 0.0, math.inf, False
        reac = []
        for t in lhs.split():
            if t == "+":
                continue
            sp, kin = parse_species(t)
            spec.setdefault(sp.name, sp)
            reac.append((sp.name, kin))
        pl = []
        for t in prods:
            sp, _ = parse_species(t)
            spec.setdefault(sp.name, sp)
            pl.append(sp.name)
        rx.append(Reaction(reac, pl, tlo, thi, ranged, rate, ln))
    # code-point order (network-format.md 2.1)
    net.species = [spec[k] for k in sorted(spec)]
    for i, r in enumerate(rx):
        r.index = i
    net.reactions = rx
    _group(net)
    _check_balance(net)
    return net


def _check_balance(net):
    for r in net.reactions:
        sp = {s.name: s for s in net.species}
        q0 = sum(sp[s].charge for s, _ in r.reactants)
        q1 = sum(sp[s].charge for s in r.products)
        if q0 != q1:
            raise KnetError("line %d: charge not conserved in %s"
                            % (r.line, r.eq))


def elements(net):
    """the elements of the network's species, sorted (the order of the
    element totals, krome_element_name)"""
    return sorted({e for sp in net.species for e in sp.comp})


def conserved_laws(net):
    """[(name, [coefficient per species])]: the nuclei of every element
    the species hold, then the charge when a species is charged, each kept
    only if every reaction conserves it"""
    elems = elements(net)
    laws = [(e, [sp.comp.get(e, 0) for sp in net.species]) for e in elems]
    if any(sp.charge for sp in net.species):
        laws.append(("charge", [sp.charge for sp in net.species]))
    idx = {sp.name: i for i, sp in enumerate(net.species)}
    kept = []
    for nm, c in laws:
        ok = all(sum(c[idx[s]] for s, _ in r.reactants)
                 == sum(c[idx[s]] for s in r.products)
                 for r in net.reactions)
        if ok:
            kept.append((nm, c))
    return kept


def jac_pattern(net, em):
    """the species block of I - gamma J as a set of (row, column): the
    diagonal; each reaction's flux rows (its reactants and products) at
    the columns of its kinetic reactants (mass action); for a rate that
    reads densities, those columns too, and its flux rows dense when it
    reads an aggregate (the reference's pattern, codegen.py:8264-8299)"""
    ns = net.nspec
    pat = {(i, i) for i in range(ns)}
    for r in net.reactions:
        rows = {net.idx(x) for x, _ in r.reactants} | {net.idx(x)
                                                       for x in r.products}
        cols = {net.idx(x) for x, kin in r.reactants if kin}
        direct, agg = em.density_deps(Expr(r.rate, net))
        cols |= direct
        if agg:
            cols = set(range(ns))
        pat |= {(i, j) for i in rows for j in cols}
    return pat


def min_degree_order(ns, pat):
    """minimum-degree elimination order of the symmetrized pattern: the
    vertex of least degree first (ties: the lowest index); eliminating it
    joins its neighbours into a clique"""
    adj = [set() for _ in range(ns)]
    for i, j in pat:
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    left = set(range(ns))
    order = []
    while left:
        v = min(left, key=lambda u: (len(adj[u]), u))
        left.remove(v)
        order.append(v)
        nb = adj[v]
        for u in nb:
            adj[u] = (adj[u] | nb) - {u, v}
        adj[v] = set()
    return order


def sparse_lu_tables(net, em):
    """the static-pivot sparse LU of the species block (the reference's
    LINSOL_SPARSE, bdf_cell.hpp:1155-1264, tables as codegen.py:8302-8341):
    perm[k] the species of pivot k; row k of L + U (fill included) at
    ci[rp[k] .. rp[k+1]) in pivot order, columns ascending; di[k] the
    position of its diagonal; flops of one factorization"""
    ns = net.nspec
    pat = jac_pattern(net, em)
    perm = min_degree_order(ns, pat)
    pos = {sp: k for k, sp in enumerate(perm)}
    rows = [set() for _ in range(ns)]
    for i, j in pat:
        rows[pos[i]].add(pos[j])
    for k in range(ns):  # symbolic elimination, no pivoting
        upper = {c for c in rows[k] if c > k}
        for i in range(k + 1, ns):
            if k in rows[i]:
                rows[i] |= upper
    rp, ci, di = [0], [], []
    for k in range(ns):
        cols = sorted(rows[k])
        di.append(len(ci) + cols.index(k))
        ci += cols
        rp.append(len(ci))
    flops = 0
    for k in range(ns):
        below = sum(1 for i in range(k + 1, ns) if k in rows[i])
        flops += below * (1 + 2 * (rp[k + 1] - di[k] - 1))
    return {"perm": perm, "rp": rp, "ci": ci, "di": di, "nnz_a": len(pat),
            "flops": flops}


# networks with more reactions than this evaluate the rhs and the species
# Jacobian by loops over index tables (the reference's RHS_LOOP_MIN,
# codegen.py:8151-8160)
RHS_LOOP_MIN = 500


def loop_tables(net):
    """index tables of the loop-form rhs and species Jacobian, in the order
    of the unrolled code (so both do the same operations in the same
    order; the reference's tables, codegen.py:8162-8223):
      kf_off/kf_sp   per reaction, its kinetic reactants (the flux factors)
      sc_off/sc_row/sc_sgn   per reaction, the rows the flux is added to
                     (reactants -1, products +1, duplicates kept)
      jt_off         per reaction, its Jacobian terms (one per distinct
                     kinetic reactant j, ascending): jt_col = j, jt_mult =
                     its multiplicity, factors jf_sp[jt_foff[t] ..] = the
                     other kinetic reactants in order"""
    T = {k: [] for k in ("kf_off", "kf_sp", "sc_off", "sc_row", "sc_sgn",
                         "jt_off", "jt_col", "jt_mult", "jt_foff", "jf_sp")}
    for r in net.reactions:
        kin = [net.idx(x) for x, k in r.reactants if k]
        T["kf_off"].append(len(T["kf_sp"]))
        T["kf_sp"] += kin
        T["sc_off"].append(len(T["sc_row"]))
        for x, _ in r.reactants:
            T["sc_row"].append(net.idx(x))
            T["sc_sgn"].append(-1)
        for x in r.products:
            T["sc_row"].append(net.idx(x))
            T["sc_sgn"].append(1)
        T["jt_off"].append(len(T["jt_col"]))
        for j in sorted(set(kin)):
            others = list(kin)
            others.remove(j)
            T["jt_col"].append(j)
            T["jt_mult"].append(kin.count(j))
            T["jt_foff"].append(len(T["jf_sp"]))
            T["jf_sp"] += others
    for off, tab in (("kf_off", "kf_sp"), ("sc_off", "sc_row"),
                     ("jt_off", "jt_col"), ("jt_foff", "jf_sp")):
        T[off].append(len(T[tab]))
    return T


# networks with more species than this are solved through strided views
# of the interleaved workspace (their vectors do not fit a lane's
# registers): the generator adds the forms below
BIG_N = 64


def sparse_positions(net, lt, sp):
    """for each (Jacobian term, scatter row) pair of the loop tables, in
    their loop order, the position of its (row, column) in the sparse LU's
    pattern (L + U in pivot order, sparse_lu_tables)"""
    pos = {e: k for k, e in enumerate(sp["perm"])}
    where = {}
    for k in range(len(sp["perm"])):
        for p in range(sp["rp"][k], sp["rp"][k + 1]):
            where[(k, sp["ci"][p])] = p
    out = []
    for r in range(len(net.reactions)):
        for t in range(lt["jt_off"][r], lt["jt_off"][r + 1]):
            col = pos[lt["jt_col"][t]]
            for q in range(lt["sc_off"][r], lt["sc_off"][r + 1]):
                key = (pos[lt["sc_row"][q]], col)
                if key not in where:
                    raise KnetError("Jacobian entry %s outside the sparse "
                                    "pattern" % (key,))
                out.append(where[key])
    return out


def _emit_big_forms(w, net, lt, sp):
    """the rhs and the analytic species Jacobian of a large network over
    any indexable state, rates and outputs (pointers, or strided views of
    the interleaved workspace); the Jacobian written straight into the
    sparse LU's pattern positions"""
    jp = sparse_positions(net, lt, sp)
    w("  // the rhs over indexable n, k, dn (same operations as rhs_loop)")
    w("  template <class NV, class KV, class DV>")
    w("  KX_HD static void rhs_loop_v(const NV& n, const KV& k, "
      "const DV& dn) {")
    w("    for (int i = 0; i < NSPEC; ++i) dn[i] = 0.0;")
    w("    for (int r = 0; r < NREA; ++r) {")
    w("      double fl = k[r];")
    w("      for (int p = LT_KF_OFF[r]; p < LT_KF_OFF[r + 1]; ++p)")
    w("        fl = fl * n[LT_KF_SP[p]];")
    w("      for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q) {")
    w("        if (LT_SC_SGN[q] > 0) dn[LT_SC_ROW[q]] += fl;")
    w("        else dn[LT_SC_ROW[q]] -= fl;")
    w("      }")
    w("    }")
    w("  }")
    w("  // the species Jacobian at fixed rates added into Jp, indexed by")
    w("  // the sparse LU's pattern position (LT_JP: one per term and")
    w("  // scatter row, in the loop order of jac_species_loop, which it")
    w("  // follows operation for operation)")
    w("  static constexpr int LT_JP[%d] = {%s};"
      % (max(len(jp), 1), ", ".join(str(v) for v in jp) or "0"))
    _emit_gather_tables(w, net, lt, jp)
    _emit_factor_columns(w, sp)
    w("  template <class NV, class KV, class JV>")
    w("  KX_HD static void jac_species_sparse(const NV& n, const KV& k, "
      "const JV& Jp) {")
    w("    int m = 0;")
    w("    for (int r = 0; r < NREA; ++r)")
    w("      for (int t = LT_JT_OFF[r]; t < LT_JT_OFF[r + 1]; ++t) {")
    w("        double d = k[r];")
    w("        for (int p = LT_JT_FOFF[t]; p < LT_JT_FOFF[t + 1]; ++p)")
    w("          d = d * n[LT_JF_SP[p]];")
    w("        if (LT_JT_MULT[t] > 1) d = d * double(LT_JT_MULT[t]);")
    w("        for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q, ++m) {")
    w("          if (LT_SC_SGN[q] > 0) Jp[LT_JP[m]] += d;")
    w("          else Jp[LT_JP[m]] -= d;")
    w("        }")
    w("      }")
    w("  }")


def gather_tables(net, lt, jp):
    """the scatter of the loop form turned around, for the kernels that
    split a cell's rhs and Jacobian over the lanes of a wave: per species
    row, its (reaction, sign) contributions in the loop form's order
    (row_off / row_r / row_sgn); per term of the Jacobian, its reaction
    (jt_r); per pattern position, its (term, sign) contributions in loop
    order (jpos_off / jpos_t / jpos_sgn).  Summed in these orders from 0,
    each sum takes the same additions in the same order as the scatter"""
    ns = net.nspec
    rows = [[] for _ in range(ns)]
    for r in range(len(net.reactions)):
        for q in range(lt["sc_off"][r], lt["sc_off"][r + 1]):
            rows[lt["sc_row"][q]].append((r, lt["sc_sgn"][q]))
    jt_r = []
    npos = max(jp) + 1 if jp else 0
    pos = [[] for _ in range(npos)]
    m = 0
    for r in range(len(net.reactions)):
        for t in range(lt["jt_off"][r], lt["jt_off"][r + 1]):
            jt_r.append(r)
            for q in range(lt["sc_off"][r], lt["sc_off"][r + 1]):
                pos[jp[m]].append((t, lt["sc_sgn"][q]))
                m += 1
    out = {"row_off": [0], "row_rs": [], "jt_r": jt_r,
           "jpos_off": [0], "jpos_ts": []}
    # a contribution's index and sign in one entry: r when added, ~r
    # (= -r - 1) when subtracted
    for lst in rows:
        out["row_rs"] += [r if g > 0 else ~r for r, g in lst]
        out["row_off"].append(len(out["row_rs"]))
    for lst in pos:
        out["jpos_ts"] += [t if g > 0 else ~t for t, g in lst]
        out["jpos_off"].append(len(out["jpos_ts"]))
    # the kinetic reactants of each reaction and the other factors of each
    # Jacobian term in fixed-width rows (-1 pads), in the loop form's order
    kw = max([lt["kf_off"][r + 1] - lt["kf_off"][r]
              for r in range(len(net.reactions))] + [1])
    out["kf_w"] = kw
    out["kfw"] = []
    for r in range(len(net.reactions)):
        sp = lt["kf_sp"][lt["kf_off"][r]:lt["kf_off"][r + 1]]
        out["kfw"] += list(sp) + [-1] * (kw - len(sp))
    nt = len(jt_r)
    jw = max([lt["jt_foff"][t + 1] - lt["jt_foff"][t] for t in range(nt)]
             + [1])
    out["jf_w"] = jw
    out["jfw"] = []
    for t in range(nt):
        sp = lt["jf_sp"][lt["jt_foff"][t]:lt["jt_foff"][t + 1]]
        out["jfw"] += list(sp) + [-1] * (jw - len(sp))
    return out


def factor_columns(sp):
    """the static-pivot factors by columns: for the strict lower part L and
    the strict upper part U, each column k's entries (row, position in the
    row-ordered factor values), rows increasing; cp[k] .. cp[k + 1] - 1"""
    ns = len(sp["di"])
    cols = {"l": [[] for _ in range(ns)], "u": [[] for _ in range(ns)]}
    for i in range(ns):
        for q in range(sp["rp"][i], sp["rp"][i + 1]):
            k = sp["ci"][q]
            if k < i:
                cols["l"][k].append((i, q))
            elif k > i:
                cols["u"][k].append((i, q))
    out = {}
    for part in ("l", "u"):
        cp, cr, cq = [0], [], []
        for lst in cols[part]:
            cr += [r for r, _ in lst]
            cq += [q for _, q in lst]
            cp.append(len(cr))
        out[part] = (cp, cr, cq)
    return out


def _emit_factor_columns(w, sp):
    fc = factor_columns(sp)
    # per factor position p (an entry of row i, column k = ci[p]): the
    # pivot position of column k and the end of U row k, so an elimination's
    # operands load without a chain through ci
    pd = [sp["di"][k] for k in sp["ci"]]
    pe = [sp["rp"][k + 1] for k in sp["ci"]]
    w("  // per factor position p (column k = sp_ci[p]): sp_di[k] and")
    w("  // sp_rp[k + 1]")
    w("  static constexpr int LT_PD[%d] = {%s};"
      % (len(pd), ", ".join(str(v) for v in pd)))
    w("  static constexpr int LT_PE[%d] = {%s};"
      % (len(pe), ", ".join(str(v) for v in pe)))
    w("  // the static-pivot factors by columns (netgen.factor_columns):")
    w("  // column k of the strict lower part L (LT_LC*) or upper part U")
    w("  // (LT_UC*) holds entries LT_?CP[k] .. LT_?CP[k + 1] - 1, each a row")
    w("  // (LT_?CR, increasing) and its position in the factor values")
    w("  // (LT_?CQ)")
    for part, tag in (("l", "LC"), ("u", "UC")):
        for nm, vals in zip(("P", "R", "Q"), fc[part]):
            vals = vals or [0]
            w("  static constexpr int LT_%s%s[%d] = {%s};"
              % (tag, nm, len(vals), ", ".join(str(v) for v in vals)))


# the lanes of a wave and the longest run of a row's contributions one
# lane sums (the wave form's rhs segments)
WAVE_LANES = 64
SEG_TERMS = 32


def row_segments(row_off, lanes=WAVE_LANES, seg=SEG_TERMS):
    """the wave form's rhs split: each species row's contributions cut into
    runs of at most seg (a row of up to seg contributions is one run), the
    runs dealt to the lanes longest first, each to the least loaded lane
    (ties: the lowest lane).  Returns (lo, hi, sid): lane l's j-th run at
    index l + lanes j covers contributions [lo, hi) and is the row-order
    run sid (empty pads: lo = hi, sid -1), and rseg_off: row i's runs are
    sid rseg_off[i] .. rseg_off[i + 1] - 1 in row order"""
    runs, rseg_off = [], [0]
    for i in range(len(row_off) - 1):
        a, b = row_off[i], row_off[i + 1]
        if b == a:
            runs.append((a, a))
        while a < b:
            runs.append((a, min(a + seg, b)))
            a += seg
        rseg_off.append(len(runs))
    load = [0] * lanes
    per = [[] for _ in range(lanes)]
    for sid in sorted(range(len(runs)),
                      key=lambda q: (-(runs[q][1] - runs[q][0]), q)):
        l = min(range(lanes), key=lambda x: (load[x], x))
        per[l].append(sid)
        load[l] += runs[sid][1] - runs[sid][0]
    depth = max(len(x) for x in per)
    lo, hi, ids = [], [], []
    for j in range(depth):
        for l in range(lanes):
            if j < len(per[l]):
                sid = per[l][j]
                lo.append(runs[sid][0])
                hi.append(runs[sid][1])
                ids.append(sid)
            else:
                lo.append(0)
                hi.append(0)
                ids.append(-1)
    return lo, hi, ids, rseg_off


def _emit_gather_tables(w, net, lt, jp):
    g = gather_tables(net, lt, jp)
    lo, hi, ids, rseg = row_segments(g["row_off"])
    w("  // the wave form's rhs runs (netgen.row_segments): lane l's j-th")
    w("  // run at l + 64 j sums contributions [LT_SEG_LO, LT_SEG_HI) of")
    w("  // LT_ROW_RS into run LT_SEG_ID (-1: a pad); row i adds its runs")
    w("  // LT_RSEG_OFF[i] .. LT_RSEG_OFF[i + 1] - 1 in order")
    for nm, vals in (("seg_lo", lo), ("seg_hi", hi), ("seg_id", ids),
                     ("rseg_off", rseg)):
        w("  static constexpr int LT_%s[%d] = {%s};"
          % (nm.upper(), len(vals), ", ".join(str(v) for v in vals)))
    w("  static constexpr int LT_NSEG = %d;" % len(lo))
    w("  static constexpr int LT_NRUN = %d;" % rseg[-1])
    w("  // the gather form of the scatter (netgen.gather_tables): per")
    w("  // species row its contributions (reaction r added, ~r subtracted),")
    w("  // per Jacobian term its reaction, per pattern position its")
    w("  // contributions (term t or ~t), each in the loop form's order; the")
    w("  // reactants of each reaction and the other factors of each term in")
    w("  // rows of LT_KF_W / LT_JF_W entries (-1 pads)")
    for nm in ("row_off", "row_rs", "jt_r", "jpos_off", "jpos_ts", "kfw",
               "jfw"):
        vals = g[nm] or [0]
        w("  static constexpr int LT_%s[%d] = {%s};"
          % (nm.upper(), len(vals), ", ".join(str(v) for v in vals)))
    w("  static constexpr int LT_NTERM = %d;" % len(g["jt_r"]))
    w("  static constexpr int LT_KF_W = %d;" % g["kf_w"])
    w("  static constexpr int LT_JF_W = %d;" % g["jf_w"])


def _group(net):
    """tile the ranged lines of identical stoichiometry into piecewise
    groups (network-format.md 3.1); an unranged line is its own group"""
    gid = 0
    by_key = {}
    for r in net.reactions:
        if not r.ranged:
            r.group = gid
            gid += 1
            continue
        key = (tuple(sorted(s for s, _ in r.reactants)),
               tuple(sorted(r.products)))
        by_key.setdefault(key, []).append(r)
    for key, rs in by_key.items():
        rs = sorted(rs, key=lambda r: r.tlo)
        chains = []
        for r in rs:
            for ch in chains:
                if ch[-1].thi == r.tlo:
                    ch.append(r)
                    break
            else:
                chains.append([r])
        for ch in chains:
            for r in ch:
                r.group = gid
            gid += 1


def groups(net):
    """the piecewise groups in the order of their first reaction"""
    out = {}
    for r in net.reactions:
        out.setdefault(r.group, []).append(r)
    return sorted((sorted(g, key=lambda r: r.tlo) for g in out.values()),
                  key=lambda g: min(r.index for r in g))


# ---------------------------------------------------------------------------
# photo reactions: photo(xsec=..., Emin=..., Emax=...)
# (docs/network-format.md 5.4; the resolution of kromex/codegen.py:
# 7530-7629 and its file readers, codegen.py:686-812)

CLIGHT = 2.99792458e10
PLANCK_EV = CONSTANTS["planck_eV"]
# h c [eV Angstrom] and [eV nm]: the SWRI files tabulate the wavelength in
# Angstrom, the Leiden files in nm (the same products as the reference)
HC_EV_AA = CLIGHT * PLANCK_EV / 1e-8
HC_EV_NM = CLIGHT * PLANCK_EV / 1e-7
# the cross-section fits a database entry may name, host C++ (kx_photo_xsec.h)
PHOTO_FITS = ("sigma_v96", "sigma_vy95")


def is_photo(r):
    return r.rate.replace(" ", "").startswith("photo(")


def _dnum(t):
    """a Fortran-style number (1.0d-3)"""
    return float(t.strip().lower().replace("d", "e"))


def photo_database(data_dir):
    """the photoionization entries of data/database/photoionization.dat:
    [{type, reacts, prods, limits, rate}] with species lists sorted and
    the electron named e-"""
    out, cur = [], None
    path = Path(data_dir) / "database" / "photoionization.dat"
    for line in path.read_text().splitlines():
        t = line.strip()
        if not t.startswith("@") or ":" not in t:
            continue
        key, val = (x.strip() for x in t[1:].split(":", 1))
        if key == "type":
            cur = {"type": val}
            out.append(cur)
        elif cur is None:
            continue
        elif key in ("reacts", "prods"):
            cur[key] = sorted("e-" if x.strip().upper() == "E" else x.strip()
                              for x in val.split(","))
        elif key == "limits":
            cur[key] = [_dnum(x) for x in val.split(",")]
        elif key == "rate":
            cur[key] = val
    return out


def leiden_xsec(data_dir, reactant, products):
    """(E [eV] ascending, sigma [cm^2]) of a Leiden database file
    <reactant>__<p1>_<p2>.dat: the photoionization column when an electron
    is among the products, else photodissociation"""
    prods = sorted(products)
    folder = Path(data_dir) / "database" / "leiden_xsecs"
    for f in sorted(folder.glob("%s__*.dat" % reactant)):
        toks = f.stem.split("__", 1)[1].split("_")
        if sorted("e-" if t.upper() == "E" else t for t in toks) == prods:
            break
    else:
        raise KnetError("no Leiden cross section for %s -> %s"
                        % (reactant, " + ".join(prods)))
    col = 3 if "e-" in prods else 2
    pts = []
    for line in f.read_text().splitlines():
        t = line.strip()
        if t and not t.startswith("#"):
            v = t.split()
            pts.append((HC_EV_NM / float(v[0]), float(v[col])))
    pts.sort()
    return [e for e, _ in pts], [x for _, x in pts]


def _swri_last_table(text):
    """(column labels, numeric rows) of the last table of an SWRI
    (phidrates) cross-section file: a table starts at its "Lambda" label
    line; lines that are not all numbers are skipped"""
    tables = []
    for line in text.splitlines():
        toks = line.split()
        if toks and toks[0] == "Lambda":
            tables.append((toks, []))
        elif toks and tables:
            try:
                tables[-1][1].append([float(x) for x in toks])
            except ValueError:
                pass
    return tables[-1] if tables else ([], [])


def _swri_products(label):
    """the products an SWRI branch label names: '/' separates them, a
    species ending in '+' brings its photoelectron, 'E' is an electron"""
    out = []
    for tok in re.findall(r"[^/+]*\+|[^/+]+", label):
        if tok.upper() == "E":
            out.append("e-")
        else:
            out.append(tok)
            if tok.endswith("+"):
                out.append("e-")
    return out


def swri_xsec(data_dir, reactant, products):
    """(E [eV] ascending, sigma [cm^2]) of the SWRI (phidrates) branch of
    <reactant>.dat whose products match: excited-state labels folded onto
    the ground state (columns summed); a run of values bracketed by the
    1e-35 sentinel on both sides is a spectral line, kept as a box of its
    mean height between the bracketing energies; other sentinels are
    dropped; the table spans its first to last nonzero value"""
    want = sorted(products)
    path = Path(data_dir) / "database" / "swri_xsecs" / ("%s.dat" % reactant)
    if not path.exists():
        raise KnetError("no SWRI file for %s" % reactant)
    label, rows = _swri_last_table(path.read_text())
    # one branch per distinct label once the excited states are folded
    # onto the ground state; its column is the sum of its labels' columns,
    # accumulated left to right
    groups = {}
    for j, h in enumerate(label[1:], 1):
        groups.setdefault(re.sub("1D|1S|3P", "", h), []).append(j)
    cols = {}
    for h, js in groups.items():
        prods = sorted(_swri_products(h))
        if prods not in [c[0] for c in cols.values()]:
            cols[h] = (prods, js)
    match = [js for prods, js in cols.values() if prods == want]
    if not match:
        raise KnetError("no SWRI branch %s in %s" % ("+".join(want),
                                                     path.name))

    def value(r, js=match[0]):
        v = r[js[0]]
        for j in js[1:]:
            v += r[j]
        return v
    pts = [(HC_EV_AA / r[0], value(r)) for r in reversed(rows) if r[0] > 0.0]
    sentinel = [v == 1e-35 for _, v in pts]
    pairs, q = [], 0
    while q < len(pts):
        if sentinel[q]:
            q += 1
            continue
        e = q
        while e < len(pts) and not sentinel[e]:
            e += 1
        run = pts[q:e]
        if q > 0 and e < len(pts):
            lo, hi = pts[q - 1][0], pts[e][0]
            h = sum(v for _, v in run) / len(run)
            pairs += [(lo * (1.0 - 1e-9), 0.0), (lo, h), (hi, h),
                      (hi * (1.0 + 1e-9), 0.0)]
        else:
            pairs += run
        q = e
    pairs.sort()
    table = []
    for e, v in pairs:
        if not table or e != table[-1][0]:
            table.append((e, v))
    while table and table[0][1] == 0.0:
        table.pop(0)
    while table and table[-1][1] == 0.0:
        table.pop()
    return [e for e, _ in table], [v for _, v in table]


# the rate database of `auto` rates (docs/network-format.md section 7):
# data/database/*.dat read as hydroHe.dat, recombination_badnell.dat, then
# the others in sorted order; the first entry whose sorted reactants and
# products match wins.  Its expression is in Tgas and the legacy
# shortcuts below, with Fortran-style numbers.
AUTO_FIRST = ("hydroHe.dat", "recombination_badnell.dat")
AUTO_SHORTCUTS = {"T32": "(Tgas/300.0)", "invT": "(1.0/Tgas)",
                  "Te": "Tgas_eV", "invTe": "(1.0/Tgas_eV)",
                  "lnTe": "log(Tgas_eV)", "T": "Tgas",
                  "sqrTgas": "sqrt(Tgas)"}
_AUTO_DB = {}


def auto_database(data_dir):
    """[(file, {type, reacts, prods, limits?, rate})] of the rate entries
    (not photoionization), in the resolution order"""
    key = str(data_dir)
    if key in _AUTO_DB:
        return _AUTO_DB[key]
    folder = Path(data_dir) / "database"
    files = [folder / f for f in AUTO_FIRST]
    files += sorted(f for f in folder.glob("*.dat")
                    if f.name not in AUTO_FIRST
                    and f.name != "photoionization.dat")
    out = []
    for f in files:
        cur = None
        for line in f.read_text(errors="replace").splitlines():
            t = line.strip()
            if not t.startswith("@") or ":" not in t:
                continue
            k, v = (x.strip() for x in t[1:].split(":", 1))
            if k == "type":
                cur = {"type": v}
                out.append((f.name, cur))
            elif cur is None:
                continue
            elif k in ("reacts", "prods"):
                cur[k] = sorted("e-" if x.strip().upper() == "E"
                                else x.strip() for x in v.split(","))
            elif k == "limits":
                parts = [re.sub(r"^(\.GE\.|\.LE\.|\.GT\.|\.LT\.|>=|<=|>|<)",
                                "", x.strip(), flags=re.I)
                         for x in v.split(",")]
                parts = ["" if x.upper() == "NONE" else x for x in parts]
                if len(parts) == 2 and any(parts):
                    cur["limits"] = (_dnum(parts[0]) if parts[0] else 0.0,
                                     _dnum(parts[1]) if parts[1]
                                     else math.inf)
            elif k == "rate":
                cur[k] = v
    _AUTO_DB[key] = out
    return out


def _auto_expr(rate):
    """a database rate in the .knet rate language"""
    expr = re.sub(r"(\d\.?\d*)[dD]([+-]?\d+)", r"\1e\2", rate)
    return re.sub(r"[A-Za-z_][A-Za-z0-9_]*",
                  lambda m: AUTO_SHORTCUTS.get(m.group(0), m.group(0)), expr)


def resolve_auto(r, data_dir):
    """the reactions an `auto` reaction becomes (codegen.py:2005-2071): the
    first matching database entry without temperature limits, the rate
    alone; or the matching entries with limits before it, each a piece
    [lo, hi) of one piecewise rate (zero below, clamped above); an
    unlimited entry after limited ones is an overlap and an error"""
    reacts = sorted(s for s, _ in r.reactants)
    prods = sorted(r.products)
    pieces = []
    for fname, e in auto_database(data_dir):
        if e.get("reacts") != reacts or e.get("prods") != prods \
                or "rate" not in e:
            continue
        if "limits" not in e:
            if pieces:
                raise KnetError("line %d: database has limited and "
                                "unlimited entries for %s" % (r.line, r.eq))
            r.rate = _auto_expr(e["rate"])
            return [r]
        pieces.append(e)
    if not pieces:
        raise KnetError("line %d: no database rate for %s" % (r.line, r.eq))
    return [Reaction(list(r.reactants), list(r.products), e["limits"][0],
                     e["limits"][1], True, _auto_expr(e["rate"]), r.line)
            for e in pieces]


def photo_reactions(net, data_dir):
    """[{r, reactant, Emin, Emax, kind, code | table}] of the photo()
    reactions in reaction order: kind "fit" (photo(xsec=auto): the
    database's outer-shell fits, piecewise from each segment's threshold),
    "table" (xsec="swri" / "leiden"), or "expr" (a function of E_eV,
    extern cross sections included)"""
    db = None
    out = []
    for r in net.reactions:
        if not is_photo(r):
            continue
        call = ast.parse(r.rate.strip(), mode="eval").body
        kw = {k.arg: k.value for k in call.keywords}
        if "xsec" not in kw or call.args:
            raise KnetError("line %d: photo() needs xsec=" % r.line)
        kin = [s for s, k in r.reactants if k]
        if len(kin) != 1:
            raise KnetError("line %d: a photo reaction has one reactant"
                            % r.line)
        ent = {"r": r.index, "reactant": net.idx(kin[0]),
               "Emin": None, "Emax": None}
        for b in ("Emin", "Emax"):
            if b in kw:
                ent[b] = float(ast.literal_eval(kw[b]))
        x = kw["xsec"]
        if isinstance(x, ast.Name) and x.id == "auto":
            if db is None:
                db = photo_database(data_dir)
            segs = sorted((e for e in db if e.get("type") == "photoion"
                           and e.get("reacts") == sorted(kin)
                           and e.get("prods") == sorted(r.products)),
                          key=lambda e: e["limits"][0])
            if not segs:
                raise KnetError("line %d: no database cross section for %s"
                                % (r.line, r.eq))
            if sum(1 for p_ in r.products if p_ == "e-") > 1:
                raise KnetError("line %d: Auger channels (several electrons) "
                                "of photo(xsec=auto) are not supported"
                                % r.line)
            code = None
            for e in segs:
                expr = re.sub(r"(\d\.?\d*)[dD]([+-]?\d+)", r"\1e\2",
                              e["rate"])
                expr = re.sub(r"\benergy_ev\b", "E_eV", expr, flags=re.I)
                names = set(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", expr))
                names -= {"E_eV", "e"} | set(PHOTO_FITS)
                if names:
                    raise KnetError("line %d: database fit %r is not "
                                    "supported" % (r.line, e["rate"]))
                expr = re.sub(r"\b(%s)\s*\(" % "|".join(PHOTO_FITS),
                              r"kx::\1(", expr)
                code = expr if code is None else "(E_eV >= %r ? %s : %s)" % (
                    e["limits"][0], expr, code)
            ent.update(kind="fit", code=code,
                       inner_shells="sigma_vy95" in code)
            if ent["Emin"] is None:
                ent["Emin"] = segs[0]["limits"][0]
        elif isinstance(x, ast.Constant) and isinstance(x.value, str):
            src = x.value.lower()
            if src == "leiden":
                tab = leiden_xsec(data_dir, kin[0], r.products)
            elif src == "swri":
                tab = swri_xsec(data_dir, kin[0], r.products)
            else:
                raise KnetError("line %d: cross-section source %r is not "
                                "supported" % (r.line, x.value))
            ent.update(kind="table", table=tab)
            if ent["Emin"] is None:
                ent["Emin"] = tab[0][0]
            if ent["Emax"] is None:
                ent["Emax"] = tab[0][-1]
        else:
            ent.update(kind="expr", ast=x)
        if ent["Emin"] is None:
            ent["Emin"] = 0.0
        if ent["Emax"] is None:
            ent["Emax"] = math.inf
        out.append(ent)
    return out


# ---------------------------------------------------------------------------
# expressions

_N_CALL = re.compile(r"\bn\(\s*([^()\s]+)\s*\)")


class Expr:
    """one rate / var expression, parsed with Python's grammar (the rate
    language is a subset of it once `n(X)` species literals are named)"""

    def __init__(self, text, net):
        self.text = text
        self.net = net
        nsub = {}

        def repl(m):
            i = net.idx(m.group(1))
            nsub["__n%d" % i] = i
            return "__n%d" % i
        t = _N_CALL.sub(repl, text)
        try:
            self.tree = ast.parse(t, mode="eval").body
        except SyntaxError as exc:
            raise KnetError("bad expression %r: %s" % (text, exc))
        self.names = {n.id for n in ast.walk(self.tree)
                      if isinstance(n, ast.Name)}
        self.calls = {n.func.id for n in ast.walk(self.tree)
                      if isinstance(n, ast.Call)}


class Emitter:
    """C++ for expressions of one network: every symbol resolves to a var
    (evaluated in the enclosing block), a state symbol, a parameter, a
    constant or a species density"""

    def __init__(self, net, param_names):
        self.net = net
        self.params = param_names
        self.exprs = {v: Expr(t, net) for v, t in net.vars.items()}

    def deps(self, expr):
        """the vars an expression needs, transitively, in dependency order"""
        out, seen = [], set()

        def visit(e, stack):
            for nm in sorted(e.names):
                if nm in self.exprs and nm not in seen:
                    if nm in stack:
                        raise KnetError("cyclic var %s" % nm)
                    visit(self.exprs[nm], stack | {nm})
                    seen.add(nm)
                    out.append(nm)
        visit(expr, set())
        return out

    def density_dependent(self, expr):
        """whether an expression (with its vars) reads the densities"""
        es = [expr] + [self.exprs[v] for v in self.deps(expr)]
        for e in es:
            if e.names & {"ntot", "Hnuclei"}:
                return True
            if any(nm.startswith("__n") for nm in e.names):
                return True
            if any(EXTERNS[c][0] for c in e.calls if c in EXTERNS):
                return True
        return False

    def density_deps(self, expr):
        """(species read directly, whether an aggregate is read): the
        densities an expression (with its vars) depends on; an aggregate
        (ntot, Hnuclei, or an extern rate function of Hnuclei) couples it
        to every species"""
        es = [expr] + [self.exprs[v] for v in self.deps(expr)]
        direct, agg = set(), False
        for e in es:
            if e.names & {"ntot", "Hnuclei"}:
                agg = True
            for c in e.calls:
                dep = EXTERNS[c][0] if c in EXTERNS else None
                if dep == "agg":
                    agg = True
                elif dep is not None:
                    direct.add(self.net.idx(dep))
            direct |= {int(nm[3:]) for nm in e.names if nm.startswith("__n")}
        return direct, agg

    def c(self, node):
        """C++ of an AST node (type R)"""
        if isinstance(node, ast.Constant):
            return "R(%r)" % float(node.value)
        if isinstance(node, ast.Name):
            nm = node.id
            if nm in self.exprs:
                return "v_" + nm
            if nm == "Tgas":
                return "Tgas"
            if nm == "E_eV":
                return "E_eV"
            if nm == "Tgas_eV":
                return "(Tgas * R(%r))" % BOLTZMANN_EV
            if nm == "ntot":
                return "ntot(n)"
            if nm == "Hnuclei":
                return "Hnuclei(n)"
            if nm.startswith("__n"):
                return "n[%d]" % int(nm[3:])
            if nm in self.params:
                return "R(p.%s)" % nm
            if nm in CONSTANTS:
                return "R(%r)" % CONSTANTS[nm]
            raise KnetError("unknown symbol %r" % nm)
        if isinstance(node, ast.UnaryOp):
            if isinstance(node.op, ast.USub):
                return "(-%s)" % self.c(node.operand)
            if isinstance(node.op, ast.UAdd):
                return self.c(node.operand)
        if isinstance(node, ast.BinOp):
            a = self.c(node.left)
            if isinstance(node.op, ast.Pow):
                ex = node.right
                val = None
                if isinstance(ex, ast.Constant):
                    val = float(ex.value)
                elif (isinstance(ex, ast.UnaryOp)
                      and isinstance(ex.op, ast.USub)
                      and isinstance(ex.operand, ast.Constant)):
                    val = -float(ex.operand.value)
                if val is not None and val == int(val) and abs(val) <= 16:
                    return "kx::ipow<%d>(%s)" % (int(val), a)
                if val is not None and val == 0.5:
                    return "kx::sqrt(%s)" % a
                return "kx::pow(%s, %s)" % (a, self.c(ex))
            b = self.c(node.right)
            op = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*",
                  ast.Div: "/"}[type(node.op)]
            return "(%s %s %s)" % (a, op, b)
        if isinstance(node, ast.Call):
            f = node.func.id
            args = node.args
            if f in ("exp", "log", "log10", "sqrt", "tanh"):
                return "kx::%s(%s)" % (f, self.c(args[0]))
            if f == "abs":
                return "kx::fabs(%s)" % self.c(args[0])
            if f in ("min", "max"):
                return "kx::f%s(%s, %s)" % (f, self.c(args[0]),
                                           self.c(args[1]))
            if f == "where":
                cnd = args[0]
                if not isinstance(cnd, ast.Compare):
                    raise KnetError("where() needs a comparison")
                op = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">",
                      ast.GtE: ">="}[type(cnd.ops[0])]
                return "((%s %s %s) ? %s : %s)" % (
                    self.c(cnd.left), op, self.c(cnd.comparators[0]),
                    self.c(args[1]), self.c(args[2]))
            if f in EXTERNS and EXTERNS[f][2] is not None:
                names = {s.name for s in self.net.species}
                return EXTERNS[f][2].format(
                    *[self.c(a) for a in args],
                    e=self.net.idx("e-") if "e-" in names else -1,
                    h2=self.net.idx("H2") if "H2" in names else -1)
            raise KnetError("unknown function %r" % f)
        raise KnetError("unsupported syntax %s" % ast.dump(node))

    def block(self, exprs, indent):
        """the var definitions the expressions need, as C++ lines"""
        need = []
        for e in exprs:
            for v in self.deps(e):
                if v not in need:
                    need.append(v)
        # dependency order over the union
        order = []
        for v in need:
            for d in self.deps(self.exprs[v]) + [v]:
                if d not in order:
                    order.append(d)
        return ["%sconst R v_%s = %s;" % (indent, v, self.c(self.exprs[v].tree))
                for v in order]


# ---------------------------------------------------------------------------
# thermal configuration

# H2 chemistry energy ledger [eV]: (reactants, products, energy, thermalized)
# (codegen.py:4953-4961)
CHEM_LEDGER = [
    (["H", "H", "H"], ["H", "H2"], 4.48, True),
    (["H", "H", "H2"], ["H2", "H2"], 4.48, True),
    (["H", "H-"], ["H2", "e-"], 3.53, True),
    (["H", "H2+"], ["H+", "H2"], 1.83, True),
    (["H", "H2"], ["H", "H", "H"], -4.48, False),
    (["H2", "e-"], ["H", "H", "e-"], -4.48, False),
    (["H2", "H2"], ["H", "H", "H2"], -4.48, False),
]
# diatomic constants [K] for thermal.gamma = EXACT: vibrational and
# rotational wavenumbers [cm^-1] times hc/k (codegen.py:318-325)
HC_K = 1.4387768775039336
GAMMA_VIB_K = {s: v * HC_K for s, v in {
    "H2": 4401.213, "HD": 3813.15, "D2": 3115.5, "CO": 2169.75589,
    "OH": 3737.761, "N2": 2358.57, "O2": 1580.161}.items()}
GAMMA_ROT_K = {s: v * HC_K for s, v in {
    "H2": 60.853, "H2+": 42.9, "HD": 45.644, "D2": 30.443, "N2": 1.9982,
    "O2": 1.4264, "CO": 2.78}.items()}

COOLING = ("H2", "CONT", "CHEM")
HEATING = ("COMPRESS", "CHEM", "PHOTO")


@dataclass
class Thermal:
    cooling: tuple = ()
    heating: tuple = ()
    gamma: str = "CONST"
    h2_opacity: str = ""
    # Tgas may be evolved even with no cooling or heating term (dT/dt = 0,
    # as the reference's thermal ODE gives with empty lists)
    evolve: bool = False

    def __post_init__(self):
        for c in self.cooling:
            if c not in COOLING:
                raise KnetError("cooling %s is not supported" % c)
        for h in self.heating:
            if h not in HEATING:
                raise KnetError("heating %s is not supported" % h)
        if self.gamma not in ("CONST", "EXACT"):
            raise KnetError("gamma %s is not supported" % self.gamma)
        if self.h2_opacity not in ("", "RIPAMONTI"):
            raise KnetError("h2_opacity %s is not supported" % self.h2_opacity)


def param_list(net, th):
    """the parameter rows' order (codegen.py:3408-3420): the network's own
    params sorted, then the predefined ones in use in their fixed order:
    gamma_ad (CONST gamma), Tcmb (always), tff (COMPRESS), opratio (EXACT
    gamma with H2), those the externs read (G0, d2g for the dust
    recombination fits) and those a rate or var names"""
    used = {"Tcmb"}
    if th.gamma == "CONST":
        used.add("gamma_ad")
    if "COMPRESS" in th.heating:
        used.add("tff")
    names = {s.name for s in net.species}
    if th.gamma == "EXACT" and "H2" in names:
        used.add("opratio")
    for e in net.externs:
        used |= set(EXTERNS[e][1])
    words = set()
    for t in list(net.vars.values()) + [r.rate for r in net.reactions]:
        words |= set(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", t))
    used |= words & set(UNIVERSAL_PARAMS)
    out = sorted(net.params)
    out += [u for u in UNIVERSAL_PARAMS if u in used and u not in out]
    return out


# ---------------------------------------------------------------------------
# the header

def _num(x):
    if math.isinf(x):
        return "INFINITY"
    if math.isnan(x):
        return "NAN"
    return repr(float(x))


def generate(knet_path, thermal=None, name=None, data_dir=None):
    """C++ header text of a network (+ its thermal configuration) and a
    metadata dict (species, masses, params) for the host side; data_dir
    holds the cross-section database and files of photo() reactions
    (default: the data/ folder beside the network's folder)"""
    net = parse_knet(knet_path)
    if data_dir is None:
        data_dir = Path(knet_path).resolve().parent.parent / "data"
    if any(r.rate.strip() == "auto" for r in net.reactions):
        rx = []
        for r in net.reactions:
            rx += resolve_auto(r, data_dir) if r.rate.strip() == "auto" \
                else [r]
        for i, r in enumerate(rx):
            r.index = i
        net.reactions = rx
        _group(net)
    photo = photo_reactions(net, data_dir)
    photo_of = {e["r"]: q for q, e in enumerate(photo)}
    NPH = len(photo)
    loop_only = net.nrea > RHS_LOOP_MIN
    th = thermal or Thermal()
    pnames = param_list(net, th)
    em = Emitter(net, pnames)
    ns = name or net.name
    NS, NR = net.nspec, net.nrea
    L = []
    w = L.append
    w("// generated by netgen.py from %s -- do not edit" % Path(knet_path).name)
    w("#pragma once")
    w('#include "kx_physics.h"')
    w("")
    w("namespace kx {")
    w("struct net_%s {" % ns)
    w('  static constexpr const char* NAME = "%s";' % ns)
    w("  static constexpr int NSPEC = %d;" % NS)
    w("  static constexpr int NREA = %d;" % NR)
    w("  static constexpr int NPAR = %d;" % len(pnames))
    for i, s in enumerate(net.species):
        w("  static constexpr int IDX_%s = %d;" % (s.cid, i))
    has_e = any(s.name == "e-" for s in net.species)
    w("  static constexpr bool HAS_E = %s;" % ("true" if has_e else "false"))
    w("  static constexpr int I_E = %d;" % (net.idx("e-") if has_e else -1))
    thermal_on = bool(th.cooling or th.heating or th.evolve)
    w("  static constexpr bool THERMAL = %s;"
      % ("true" if thermal_on else "false"))
    hnames, cnames = thermal_term_names(th)
    w("  static constexpr int NHEAT = %d;" % len(hnames))
    w("  static constexpr int NCOOL = %d;" % len(cnames))
    # n-dependent rates -> no analytic Jacobian
    grp = groups(net)
    dens_groups = [g for g in grp
                   if any(not is_photo(r)
                          and em.density_dependent(Expr(r.rate, net))
                          for r in g)]
    w("  static constexpr bool RATES_DENSITY = %s;"
      % ("true" if dens_groups else "false"))
    w("  // photo reactions (photo(xsec=...)): their rates are the bin sums")
    w("  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read")
    w("  // after the parameters, in the row's last NPHOTO slots")
    photo_heat = "PHOTO" in th.heating
    if photo_heat and not NPH:
        raise KnetError("PHOTO heating without photo reactions")
    if photo_heat and any(e.get("inner_shells") for e in photo):
        raise KnetError("PHOTO heating with inner-shell photoionization "
                        "fits (their per-shell deposits) is not supported")
    w("  static constexpr int NPHOTO = %d;" % NPH)
    w("  // PHOTO heating: the energy sums sum_b sigmaE_qb Phi_b follow the")
    w("  // rate sums in the row (sigmaE: sigma times the photoelectron's")
    w("  // energy E - Emin_q, bin-averaged)")
    w("  static constexpr bool PHOTO_HEAT = %s;"
      % ("true" if photo_heat else "false"))
    w("  static constexpr int NROW = NPAR + NPHOTO%s;"
      % (" * 2" if photo_heat else ""))
    w("  // parameter row (param_names order), then the photo sums")
    w("  struct Params {")
    for p in pnames:
        w("    double %s;" % p)
    if NPH:
        w("    double ph[NPHOTO];")
    if photo_heat:
        w("    double phe[NPHOTO];")
    w("  };")
    w("  KX_HD static Params params_from_row(const double* row) {")
    w("    Params p;")
    for i, p in enumerate(pnames):
        w("    p.%s = row[%d];" % (p, i))
    if NPH:
        w("    for (int r = 0; r < NPHOTO; ++r) p.ph[r] = row[NPAR + r];")
    if photo_heat:
        w("    for (int r = 0; r < NPHOTO; ++r)")
        w("      p.phe[r] = row[NPAR + NPHOTO + r];")
    w("    return p;")
    w("  }")
    _emit_photo(w, net, em, photo)
    req = [i for i, p in enumerate(pnames)
           if math.isnan(net.params.get(p, PARAM_DEFAULTS.get(p, 0.0)))]
    w("  // index of the first REQUIRED parameter still NaN, or -1")
    w("  KX_HD static int params_missing(const double* row) {")
    for i in req:
        w("    if (row[%d] != row[%d]) return %d;" % (i, i, i))
    w("    (void)row;")
    w("    return -1;")
    w("  }")
    # species data
    charges = [s.charge for s in net.species]
    masses = [s.mass_amu * AMU_G for s in net.species]
    natH = [float(s.comp.get("H", 0)) for s in net.species]
    w("  KX_HD static int charge(int i) {")
    w("    switch (i) {")
    for i, q in enumerate(charges):
        if q:
            w("      case %d: return %d;" % (i, q))
    w("      default: return 0;")
    w("    }")
    w("  }")
    w("  KX_HD static double mass_g(int i) {")
    w("    switch (i) {")
    for i, m in enumerate(masses):
        w("      case %d: return %r;" % (i, m))
    w("      default: return 0.0;")
    w("    }")
    w("  }")
    # element composition (query.hpp:152-170, codegen.py:3544-3566): the
    # elements the species hold, sorted; natoms(i, e) atoms of element e in
    # species i (every species here is an element composition: none opaque)
    elems = elements(net)
    w("  static constexpr int NELEM = %d;" % len(elems))
    w("  KX_HD static constexpr int natoms(int i, int e) {")
    w("    switch (i * %d + e) {" % max(len(elems), 1))
    for i, sp in enumerate(net.species):
        for e, el in enumerate(elems):
            if sp.comp.get(el, 0):
                w("      case %d: return %d;" % (i * len(elems) + e,
                                               sp.comp[el]))
    w("      default: return 0;")
    w("    }")
    w("  }")
    # the static-pivot sparse LU of the species block (linsol 3)
    sp = sparse_lu_tables(net, em)
    w("  // static-pivot sparse LU of the species block of I - gamma J:")
    w("  // %d nonzeros, %d in L + U with the fill, %d flops per"
      % (sp["nnz_a"], len(sp["ci"]), sp["flops"]))
    w("  // factorization (dense: %d); rows in pivot order" % (2 * NS ** 3 // 3))
    w("  static constexpr int SP_NNZ = %d;" % len(sp["ci"]))
    w("  static constexpr int SP_FLOPS = %d;" % sp["flops"])
    for nm in ("perm", "rp", "ci", "di"):
        w("  static constexpr int sp_%s[%d] = {%s};"
          % (nm, len(sp[nm]), ", ".join(str(v) for v in sp[nm])))
    # conserved combinations: the nuclei of each element and the charge,
    # kept where every reaction keeps them
    laws = conserved_laws(net)
    w("  // conserved combinations c_k . n (the nuclei of %s%s): every"
      % (", ".join(nm for nm, _ in laws if nm != "charge"),
         ", the charge" if any(nm == "charge" for nm, _ in laws) else ""))
    w("  // reaction keeps them, so c_k^T J = 0 over the species")
    w("  static constexpr int NCONS = %d;" % len(laws))
    w("  KX_HD static constexpr int cons(int k, int i) {")
    w("    switch (k * NSPEC + i) {")
    for k, (_, c) in enumerate(laws):
        for i, v in enumerate(c):
            if v:
                w("      case %d: return %d;" % (k * NS + i, v))
    w("      default: return 0;")
    w("    }")
    w("  }")
    w("  template <class R> KX_HD static R ntot(const R* n) {")
    w("    R s = R(0);")
    for i in range(NS):
        w("    s += n[%d];" % i)
    w("    return s;")
    w("  }")
    w("  template <class R> KX_HD static R Hnuclei(const R* n) {")
    w("    R s = R(0);")
    for i in range(NS):
        if natH[i]:
            w("    s += n[%d] * R(%r);" % (i, natH[i]))
        else:
            w("    s += n[%d] * R(0.0);" % i)
    w("    return s;")
    w("  }")
    w("  template <class R> KX_HD static R rho(const R* n) {")
    w("    R s = R(0);")
    for i in range(NS):
        w("    s += n[%d] * R(%r);" % (i, masses[i]))
    w("    return s;")
    w("  }")

    # rates
    def emit_group(g, indent):
        out = []
        if len(g) == 1 and g[0].index in photo_of:
            r = g[0]
            return ["%sk[%d] = R(p.ph[%d]);  // %s (photo)"
                    % (indent, r.index, photo_of[r.index], r.eq)]
        ex = [Expr(r.rate, net) for r in g]
        if not g[0].ranged:
            r = g[0]
            blk = em.block(ex, indent + "  ")
            line = "k[%d] = %s;  // %s" % (r.index, em.c(ex[0].tree), r.eq)
            if not blk:
                return [indent + line]
            return ([indent + "{"] + blk + [indent + "  " + line]
                    + [indent + "}"])
        lo, hi = g[0].tlo, g[-1].thi
        out.append("%s{  // %s: %d piece(s), zero below %s, clamped above %s"
                   % (indent, g[0].eq, len(g), _num(lo), _num(hi)))
        ind = indent + "  "
        out.append("%sR Tgas = Trate;" % ind)
        if not math.isinf(hi):
            out.append("%sif (Tgas > R(%r)) Tgas = R(%r);" % (ind, hi, hi))
        for ln in em.block(ex, ind):
            out.append(ln)
        for j, (r, e) in enumerate(zip(g, ex)):
            conds = []
            if r.tlo > 0.0:
                conds.append("Tgas >= R(%r)" % r.tlo)
            if j < len(g) - 1:
                conds.append("Tgas < R(%r)" % r.thi)
            cnd = " && ".join(conds) or "true"
            out.append("%sif (%s) k[%d] = %s;  // %s" % (
                ind, cnd, r.index, em.c(e.tree), r.eq))
        out.append("%s}" % indent)
        return out

    big = net.nspec > BIG_N
    w("  // rates of the rate temperature Trate (and parameters); the")
    w("  // caller zeroes k%s" % (" (any indexable k: a pointer, or a view of"
                                " an interleaved workspace)" if big else ""))
    if big:
        w("  template <class R, class KV>")
        w("  KX_HD static void rates_T(R Trate, const Params& p, KV k) {")
    else:
        w("  template <class R>")
        w("  KX_HD static void rates_T(R Trate, const Params& p, R* k) {")
    w("    (void)p;")
    w("    const R Tgas = Trate;")
    w("    (void)Tgas;")
    tgroups = [g for g in grp if g not in dens_groups]
    for g in tgroups:
        L.extend(emit_group(g, "    "))
    w("  }")
    w("  // density-dependent rates")
    w("  template <class R>")
    w("  KX_HD static void rates_n(R Trate, const R* n, const Params& p, R* k) {")
    w("    (void)p; (void)n;")
    w("    const R Tgas = Trate;")
    w("    (void)Tgas;")
    for g in dens_groups:
        L.extend(emit_group(g, "    "))
    w("  }")
    w("  // every reaction: the T-only ones in k (zeroed) then the rest")
    w("  template <class R>")
    w("  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {")
    w("    for (int i = 0; i < NREA; ++i) k[i] = R(0);")
    w("    rates_T(Trate, p, k);")
    w("    rates_n(Trate, n, p, k);")
    w("  }")
    w("  static constexpr int NREA_T = %d;"
      % sum(len(g) for g in tgroups))

    # rhs and species Jacobian: unrolled (up to RHS_LOOP_MIN reactions),
    # and as loops over index tables (always; the only form above it)
    lt = loop_tables(net)
    w("  // the loop form's index tables (netgen.loop_tables)")
    w("  static constexpr bool RHS_LOOP = %s;"
      % ("true" if loop_only else "false"))
    for nm in ("kf_off", "kf_sp", "sc_off", "sc_row", "sc_sgn", "jt_off",
               "jt_col", "jt_mult", "jt_foff", "jf_sp"):
        vals = lt[nm] or [0]
        w("  static constexpr int LT_%s[%d] = {%s};"
          % (nm.upper(), len(vals), ", ".join(str(v) for v in vals)))
    w("  // the reaction fluxes k prod(n) [cm^-3 s^-1] (krome_get_fluxes)")
    w("  template <class R>")
    w("  KX_HD static void fluxes(const R* n, const R* k, R* fl) {")
    w("    for (int r = 0; r < NREA; ++r) {")
    w("      R f = k[r];")
    w("      for (int p = LT_KF_OFF[r]; p < LT_KF_OFF[r + 1]; ++p)")
    w("        f = f * n[LT_KF_SP[p]];")
    w("      fl[r] = f;")
    w("    }")
    w("  }")
    w("  // dn/dt = sum over reactions of S k prod(n), in reaction order")
    w("  template <class R>")
    w("  KX_HD static void rhs_loop(const R* n, const R* k, R* dn) {")
    w("    for (int i = 0; i < NSPEC; ++i) dn[i] = R(0);")
    w("    for (int r = 0; r < NREA; ++r) {")
    w("      R fl = k[r];")
    w("      for (int p = LT_KF_OFF[r]; p < LT_KF_OFF[r + 1]; ++p)")
    w("        fl = fl * n[LT_KF_SP[p]];")
    w("      for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q) {")
    w("        if (LT_SC_SGN[q] > 0) dn[LT_SC_ROW[q]] += fl;")
    w("        else dn[LT_SC_ROW[q]] -= fl;")
    w("      }")
    w("    }")
    w("  }")
    w("  // d(dn_i)/d(n_j) at fixed rates into J (column-major, element")
    w("  // (i, j) at J[i + j ld]; added to, the caller zeroes it)")
    w("  template <class R>")
    w("  KX_HD static void jac_species_loop(const R* n, const R* k, R* J,")
    w("                                     int ld) {")
    w("    for (int r = 0; r < NREA; ++r)")
    w("      for (int t = LT_JT_OFF[r]; t < LT_JT_OFF[r + 1]; ++t) {")
    w("        R d = k[r];")
    w("        for (int p = LT_JT_FOFF[t]; p < LT_JT_FOFF[t + 1]; ++p)")
    w("          d = d * n[LT_JF_SP[p]];")
    w("        if (LT_JT_MULT[t] > 1) d = d * R(LT_JT_MULT[t]);")
    w("        R* col = J + (long)LT_JT_COL[t] * ld;")
    w("        for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q) {")
    w("          if (LT_SC_SGN[q] > 0) col[LT_SC_ROW[q]] += d;")
    w("          else col[LT_SC_ROW[q]] -= d;")
    w("        }")
    w("      }")
    w("  }")
    if big:
        _emit_big_forms(w, net, lt, sp)
    if not loop_only:
        w("  template <class R>")
        w("  KX_HD static void rhs_unrolled(const R* n, const R* k, R* dn) {")
        w("    for (int i = 0; i < NSPEC; ++i) dn[i] = R(0);")
        w("    R fl;")
        for r in net.reactions:
            prod = "*".join(["k[%d]" % r.index] + ["n[%d]" % net.idx(s)
                                                    for s, kin in r.reactants
                                                    if kin])
            w("    fl = %s;  // %s" % (prod, r.eq))
            for s, _ in r.reactants:
                w("    dn[%d] -= fl;" % net.idx(s))
            for s in r.products:
                w("    dn[%d] += fl;" % net.idx(s))
        w("    (void)fl;")
        w("  }")
        w("  template <class R>")
        w("  KX_HD static void jac_species_unrolled(const R* n, const R* k, "
          "R* J,")
        w("                                         int ld) {")
        w("    R d;")
        for r in net.reactions:
            kin = [net.idx(s) for s, k_ in r.reactants if k_]
            for j in sorted(set(kin)):
                mult = kin.count(j)
                others = list(kin)
                others.remove(j)
                terms = ["k[%d]" % r.index] + ["n[%d]" % o for o in others]
                if mult > 1:
                    terms.append("R(%r)" % float(mult))
                w("    d = %s;  // %s, d/d n[%d]" % ("*".join(terms), r.eq,
                                                   j))
                for s, _ in r.reactants:
                    w("    J[%d + %d * ld] -= d;" % (net.idx(s), j))
                for s in r.products:
                    w("    J[%d + %d * ld] += d;" % (net.idx(s), j))
        w("    (void)d; (void)n; (void)k; (void)J; (void)ld;")
        w("  }")
    # the form the solvers call: the loop form above RHS_LOOP_MIN
    # reactions, or everywhere when compiled with KX_FORCE_RHS_LOOP
    w("#if defined(KX_FORCE_RHS_LOOP)")
    w("  static constexpr bool RHS_LOOP_USED = true;")
    w("#else")
    w("  static constexpr bool RHS_LOOP_USED = RHS_LOOP;")
    w("#endif")
    w("  template <class R>")
    w("  KX_HD static void rhs(const R* n, const R* k, R* dn) {")
    if loop_only:
        w("    rhs_loop(n, k, dn);")
    else:
        w("    if constexpr (RHS_LOOP_USED) rhs_loop(n, k, dn);")
        w("    else rhs_unrolled(n, k, dn);")
    w("  }")
    w("  template <class R>")
    w("  KX_HD static void jac_species(const R* n, const R* k, R* J, int ld) {")
    if loop_only:
        w("    jac_species_loop(n, k, J, ld);")
    else:
        w("    if constexpr (RHS_LOOP_USED) jac_species_loop(n, k, J, ld);")
        w("    else jac_species_unrolled(n, k, J, ld);")
    w("  }")

    # thermal
    if thermal_on:
        _emit_thermal(w, net, th, pnames, photo)
    w("};")
    w("}  // namespace kx")
    meta = {"name": ns, "species": [s.name for s in net.species],
            "charge": charges, "mass_g": masses, "params": pnames,
            "elements": elems,
            "heating": hnames, "cooling": cnames,
            "sparse": {k: sp[k] for k in ("perm", "rp", "ci", "di",
                                          "flops")},
            "natoms": [[sp.comp.get(e, 0) for e in elems]
                       for sp in net.species],
            "param_defaults": [net.params.get(p, PARAM_DEFAULTS.get(p))
                               for p in pnames],
            "nrea": NR, "thermal": thermal_on,
            "photo": [{"reaction": e["r"], "reactant": e["reactant"],
                       "Emin": e["Emin"], "Emax": e["Emax"]}
                      for e in photo],
            "reactions": [r.eq for r in net.reactions],
            "rates_density": bool(dens_groups)}
    return "\n".join(L) + "\n", meta


def _emit_photo(w, net, em, photo):
    """the photo reactions' tables and cross sections: photo_reaction,
    photo_reactant, the energy window [photo_Emin, photo_Emax] [eV], and
    photo_xsec(q, E) [cm^2], a host function for the bins' set-up"""
    n = max(len(photo), 1)

    def arr(typ, nm, vals, fmt):
        vals = list(vals) or [0]
        w("  static constexpr %s %s[%d] = {%s};"
          % (typ, nm, n, ", ".join(fmt(v) for v in vals)))
    arr("int", "photo_reaction", (e["r"] for e in photo), str)
    arr("int", "photo_reactant", (e["reactant"] for e in photo), str)
    arr("double", "photo_Emin", (e["Emin"] for e in photo), _num)
    arr("double", "photo_Emax", (e["Emax"] for e in photo), _num)
    for q, e in enumerate(photo):
        if e["kind"] != "table":
            continue
        E, S = e["table"]
        w("  // tabulated cross section of %s: %d points, %.4g-%.4g eV"
          % (net.reactions[e["r"]].eq, len(E), E[0], E[-1]))
        for nm, vals in (("E", E), ("S", S)):
            w("  static constexpr double PX%d_%s[%d] = {" % (q, nm, len(vals)))
            for i in range(0, len(vals), 6):
                w("    " + ", ".join(repr(float(v)) for v in vals[i:i + 6])
                  + ",")
            w("  };")
    w("  // cross section [cm^2] of photo reaction q at photon energy E_eV")
    w("  static double photo_xsec(int q, double E_eV) {")
    w("    using R = double;")
    w("    switch (q) {")
    for q, e in enumerate(photo):
        if e["kind"] == "fit":
            code = e["code"]
        elif e["kind"] == "table":
            code = "kx::xsec_interp(E_eV, PX%d_E, PX%d_S, %d)" % (
                q, q, len(e["table"][0]))
        else:
            x = e["ast"]
            if (isinstance(x, ast.Call) and isinstance(x.func, ast.Name)
                    and x.func.id in EXTERNS
                    and EXTERNS[x.func.id][2] is None):
                code = "kx::%s(%s)" % (x.func.id, ", ".join(
                    em.c(a) for a in x.args))
            else:
                code = em.c(x)
        w("      case %d: return %s;  // %s"
          % (q, code, net.reactions[e["r"]].eq))
    w("      default: return R(0) * E_eV;")
    w("    }")
    w("  }")


# the Lyman-Werner band [eV] (kromex/knet.py LW_BAND)
LW_BAND = (11.2, 13.6)


def is_solomon(r, photo_entry=None):
    """H2 -> H + H by Lyman-Werner line pumping (kromex/codegen.py:
    7210-7230): a rate through H2_solomonLW, or a photo rate whose window
    ends inside the band.  Such reactions heat by Solomon pumping (CHEM)
    and deposit no photoelectron"""
    R = sorted(s for s, _ in r.reactants)
    if R != ["H2"] or sorted(r.products) != ["H", "H"]:
        return False
    if photo_entry is not None:
        return photo_entry["Emax"] <= LW_BAND[1] + 1e-9
    return "H2_solomonLW" in re.findall(r"[A-Za-z_][A-Za-z0-9_]*", r.rate)


def _emit_thermal(w, net, th, pnames, photo=()):
    names = {s.name for s in net.species}
    idx = net.idx
    w("  // thermal terms at the thermal temperature Tth (capped to")
    w("  // [2.73, 1e9] K by the caller) with the rates k of the state")
    if "CHEM" in th.cooling or "CHEM" in th.heating:
        heat, cool = [], []
        for r in net.reactions:
            R_ = sorted(s for s, _ in r.reactants)
            P_ = sorted(r.products)
            for rr, pp, q, ff in CHEM_LEDGER:
                if R_ == rr and P_ == pp:
                    (cool if q < 0 else heat).append((r, abs(q), ff))
                    break
    if "CHEM" in th.cooling:
        w("  template <class R>")
        w("  KX_HD static R cooling_chem(const R* n, const R* k) {")
        w("    R c = R(0);")
        for r, q, _ in cool:
            prod = "*".join("n[%d]" % idx(s) for s, _ in r.reactants)
            w("    c += k[%d] * R(%r) * %s;  // %s" % (r.index, q, prod, r.eq))
        w("    return c * R(%r);" % EV_TO_ERG)
        w("  }")
    if "CHEM" in th.heating:
        # the formation heating, then Solomon pumping: (18.7 ffac + 0.4) eV
        # per dissociation (codegen.py:4952-5029)
        pe = {e["r"]: e for e in photo}
        sol = [r for r in net.reactions if is_solomon(r, pe.get(r.index))]
        w("  template <class R>")
        w("  KX_HD static R heating_chem(R Tth, const R* n, const R* k) {")
        w("    R h = R(0);")
        if heat or sol:
            w("    const R ffac = kx::h2_heat_fraction(Tth, n[%d], n[%d], "
              "Hnuclei(n));" % (idx("H"), idx("H2")))
            for r, q, ff in heat:
                prod = "*".join("n[%d]" % idx(s) for s, _ in r.reactants)
                w("    h += k[%d] * (R(%r) * ffac) * %s;  // %s"
                  % (r.index, q, prod, r.eq))
            for r in sol:
                w("    h += k[%d] * (R(18.7) * ffac + R(0.4)) * n[%d];"
                  "  // Solomon: %s" % (r.index, idx("H2"), r.eq))
        w("    (void)Tth;")
        w("    return h * R(%r);" % EV_TO_ERG)
        w("  }")
    if "PHOTO" in th.heating:
        # photo.hpp:322-342: each absorption deposits the photoelectron's
        # energy; the sum over the photo reactions, then eV -> erg
        w("  template <class R>")
        w("  KX_HD static R heating_photo(const R* n, const Params& p) {")
        w("    R h = R(0);")
        for q, e in enumerate(photo):
            if is_solomon(net.reactions[e["r"]], e):
                continue  # line pumping: no photoelectron
            w("    h += R(p.phe[%d]) * n[%d];  // %s"
              % (q, e["reactant"], net.reactions[e["r"]].eq))
        w("    return h * R(%r);" % EV_TO_ERG)
        w("  }")
    # gamma
    w("  template <class R>")
    w("  KX_HD static R gamma_index(R Tth, const R* n, const Params& p) {")
    if th.gamma == "CONST":
        w("    (void)Tth; (void)n;")
        w("    return R(p.gamma_ad);")
    else:
        mono = [s for s in net.species
                if (s.name == "e-") or sum(s.comp.values()) == 1]
        dis = [s for s in net.species if s.name != "e-"
               and sum(s.comp.values()) == 2
               and (s.name in GAMMA_VIB_K or s.name in GAMMA_ROT_K)]
        ves = [GAMMA_VIB_K[s.name] for s in dis if s.name in GAMMA_VIB_K]
        if ves:
            w("    const R iTv = R(1) / kx::fmax(Tth, R(%r));"
              % (min(ves) * 1e-2))
        terms = []
        for s in dis:
            vp = "R(0)"
            if s.name in GAMMA_VIB_K:
                w("    const R x_%s = R(%r) * iTv;"
                  % (s.cid, GAMMA_VIB_K[s.name]))
                vp = "R(2) * kx::vib_heat_capacity(x_%s)" % s.cid
            rp = "R(2)"
            if s.name == "H2":
                rp = ("R(2) * kx::gamma_rot_ortho_para(Tth, R(p.opratio) / "
                      "(R(p.opratio) + R(1)))")
            elif s.name in GAMMA_ROT_K:
                rp = "R(2) * kx::gamma_rot(Tth, R(%r))" % GAMMA_ROT_K[s.name]
            w("    const R gi_%s = R(0.5) * (R(3) + %s + %s);"
              % (s.cid, rp, vp))
            terms.append(s)
        w("    const R mosum = %s;" % (" + ".join("n[%d]" % idx(s.name)
                                              for s in mono) or "R(0)"))
        w("    const R ysum = mosum%s;"
          % "".join(" + n[%d]" % idx(s.name) for s in terms))
        w("    const R gsum = R(1.5) * mosum%s;"
          % "".join(" + n[%d] * gi_%s" % (idx(s.name), s.cid)
                    for s in terms))
        w("    (void)p;")
        w("    return gsum > R(0) ? R(1) + ysum / gsum : R(5) / R(3);")
    w("  }")
    # dT/dt
    w("  // dTgas/dt: heating minus cooling (the reference's term order:")
    w("  // heat CHEM, COMPRESS, PHOTO; cool H2, CHEM, CONT)")
    w("  template <class R>")
    w("  KX_HD static R thermal(R Tth, const R* n, const R* k, const Params& p) {")
    w("    R heat = R(0), cool = R(0);")
    if "CHEM" in th.heating:
        w("    heat += heating_chem(Tth, n, k);")
    if "COMPRESS" in th.heating:
        w("    heat += ntot(n) * R(%r) * Tth / R(p.tff);" % BOLTZMANN_ERG)
    if "PHOTO" in th.heating:
        w("    heat += heating_photo(n, p);")
    if "H2" in th.cooling:
        if "H2" not in names:
            raise KnetError("H2 cooling without H2")
        args = ", ".join("n[%d]" % idx(x) if x in names else "R(0)"
                         for x in ("H", "H+", "H2", "e-", "He"))
        op = ("kx::h2_opacity_ripamonti(ntot(n))"
              if th.h2_opacity == "RIPAMONTI" else "R(1)")
        w("    cool += kx::cooling_H2(Tth, %s) * %s;" % (args, op))
    if "CHEM" in th.cooling:
        w("    cool += cooling_chem(n, k);")
    if "CONT" in th.cooling:
        w("    cool += kx::cooling_continuum(Tth, rho(n), ntot(n));")
    w("    return (gamma_index(Tth, n, p) - R(1)) * (heat - cool)")
    w("        / (R(%r) * ntot(n));" % BOLTZMANN_ERG)
    w("  }")
    # the same terms one by one, for the thermal query (krome_get_thermal)
    hs, cs = thermal_term_names(th)
    w("  // the active thermal terms one by one, in the order of the names")
    w("  // (heat: %s; cool: %s), the same expressions as thermal()"
      % (", ".join(hs), ", ".join(cs)))
    w("  template <class R>")
    w("  KX_HD static void thermal_terms(R Tth, const R* n, const R* k,")
    w("                                  const Params& p, R* heats, R* cools) {")
    for i, t in enumerate(hs):
        if t == "CHEM":
            w("    heats[%d] = heating_chem(Tth, n, k);" % i)
        elif t == "COMPRESS":
            w("    heats[%d] = ntot(n) * R(%r) * Tth / R(p.tff);"
              % (i, BOLTZMANN_ERG))
        elif t == "PHOTO":
            w("    heats[%d] = heating_photo(n, p);" % i)
    for i, t in enumerate(cs):
        if t == "H2":
            args = ", ".join("n[%d]" % idx(x) if x in names else "R(0)"
                             for x in ("H", "H+", "H2", "e-", "He"))
            op = ("kx::h2_opacity_ripamonti(ntot(n))"
                  if th.h2_opacity == "RIPAMONTI" else "R(1)")
            w("    cools[%d] = kx::cooling_H2(Tth, %s) * %s;" % (i, args, op))
        elif t == "CHEM":
            w("    cools[%d] = cooling_chem(n, k);" % i)
        elif t == "CONT":
            w("    cools[%d] = kx::cooling_continuum(Tth, rho(n), ntot(n));"
              % i)
    w("    (void)Tth; (void)n; (void)k; (void)p; (void)heats; (void)cools;")
    w("  }")
    # the derivatives of heating minus cooling by each rate: the CHEM terms
    # are the only thermal terms that read k, each linear in its rate (the
    # adjoint's d(dT/dt)/d log k)
    w("  // d(heat - cool)/dk_r [erg cm^-3 s^-1 per unit rate] into c[NREA]:")
    w("  // the CHEM terms, the only thermal terms that read the rates")
    w("  template <class R>")
    w("  KX_HD static void chem_dk(R Tth, const R* n, R* c) {")
    w("    for (int r = 0; r < NREA; ++r) c[r] = R(0);")
    if "CHEM" in th.heating:
        pe = {e["r"]: e for e in photo}
        sol = [r for r in net.reactions if is_solomon(r, pe.get(r.index))]
        if heat or sol:
            w("    const R ffac = kx::h2_heat_fraction(Tth, n[%d], n[%d], "
              "Hnuclei(n));" % (idx("H"), idx("H2")))
            for r, q, ff in heat:
                prod = "*".join("n[%d]" % idx(s) for s, _ in r.reactants)
                w("    c[%d] += (R(%r) * ffac) * %s * R(%r);"
                  % (r.index, q, prod, EV_TO_ERG))
            for r in sol:
                w("    c[%d] += (R(18.7) * ffac + R(0.4)) * n[%d] * R(%r);"
                  % (r.index, idx("H2"), EV_TO_ERG))
    if "CHEM" in th.cooling:
        for r, q, _ in cool:
            prod = "*".join("n[%d]" % idx(s) for s, _ in r.reactants)
            w("    c[%d] -= R(%r) * %s * R(%r);" % (r.index, q, prod,
                                                    EV_TO_ERG))
    w("    (void)Tth; (void)n;")
    w("  }")


# the reference's fixed lists of thermal terms (solve_cell.hpp:206-231):
# a build reports its active ones in this order
HEAT_TERM_ORDER = ("CUSTOM", "CR", "CHEM", "DH", "COMPRESS", "VISCOUS",
                   "PHOTODUST", "PHOTODUSTNET", "PHOTOAV", "PHOTO", "ZCIE",
                   "NUCLEAR")
COOL_TERM_ORDER = ("CUSTOM", "ATOMIC", "H2", "HD", "MOL", "Z", "Z_CIE",
                   "Z_EXTENDED", "FLOOR", "CHEM", "COMPTON", "CIE",
                   "EXPANSION", "CONT", "FF", "DH", "DUST", "DUSTTAB",
                   "DUSTBINS")


def thermal_term_names(th):
    """(heating names, cooling names) of a thermal configuration's active
    terms, in the reference's order"""
    return ([t for t in HEAT_TERM_ORDER if t in th.heating],
            [t for t in COOL_TERM_ORDER if t in th.cooling])


def meta_header(meta, th):
    """the names and defaults a network's C ABI reports (krome_species_name,
    krome_param_name, krome_params_default, krome_param_required_by)"""
    def cstr(xs):
        return ", ".join('"%s"' % x for x in xs)
    req = []
    for p, d in zip(meta["params"], meta["param_defaults"]):
        if d is not None and isinstance(d, float) and math.isnan(d):
            req.append("thermal.heating = ['COMPRESS']" if p == "tff"
                       else "the network")
        else:
            req.append("")
    L = ["// generated by netgen.py -- do not edit",
         "// the species and parameter names of network %s" % meta["name"],
         "const char* const kSpecies[] = {%s};" % cstr(meta["species"]),
         "const char* const kParams[] = {%s};" % cstr(meta["params"]),
         "const char* const kElements[] = {%s};" % cstr(
             meta["elements"] or [""]),
         "const char* const kHeatNames[] = {%s};" % cstr(
             meta["heating"] or [""]),
         "const char* const kCoolNames[] = {%s};" % cstr(
             meta["cooling"] or [""]),
         "const double kParamDefaults[] = {%s};" % ", ".join(
             "NAN" if (isinstance(d, float) and math.isnan(d)) else repr(float(d))
             for d in meta["param_defaults"]),
         "const char* const kParamRequiredBy[] = {%s};" % cstr(req)]
    return "\n".join(L) + "\n"


# the configurations this project builds: (network file, thermal options)
CONFIGS = {
    "primordial": ("primordial.knet", Thermal()),
    "primordial3": ("primordial3.knet",
                    Thermal(cooling=("H2", "CONT", "CHEM"),
                            heating=("COMPRESS", "CHEM"), gamma="EXACT",
                            h2_opacity="RIPAMONTI")),
    # photochemistry: 8 photo() reactions over photon bins, the Solomon
    # rate and grain-assisted recombination; the default thermal options
    # (no cooling or heating term), Tgas evolved
    "galaxy_ism": ("galaxy_ism.knet", Thermal(evolve=True)),
    # 27 species with the metals C and O, density-gated rates
    # (where(ntot > 1e14, ...)) and two database rates; Tgas fixed
    "primordialZ": ("primordialZ.knet", Thermal()),
    # galaxy_ism with a thermal balance: photoheating of the fields and the
    # chemical heating against H2 and chemical cooling, Tgas evolved
    "galaxy_ism_heat": ("galaxy_ism.knet",
                        Thermal(cooling=("H2", "CHEM"),
                                heating=("CHEM", "PHOTO"))),
    # the uneven batch's third network: Herbst's molecular-cloud chemistry,
    # 452 species (GRAIN0 and GRAIN- opaque) and 4429 reactions, the loop
    # form of the rhs and the static-pivot sparse LU; Tgas fixed
    "cloud": ("cloud.knet", Thermal()),
}
# the other shipped networks this generator covers, at the default thermal
# options (Tgas fixed), each pinned against the reference's generated code
# and solver on a generic seeded workload (cells.generic_cells)
EXTRA = ("NO", "SM", "coolHCN", "lowmetal", "primordial2", "primordial_wD",
         "primordial_photo", "primordial_photoH2", "auto", "chianti", "ions",
         "ionsH2", "primordial_UV", "xrays", "primordialZ_UV", "photobin")
CONFIGS.update({n: ("%s.knet" % n, Thermal()) for n in EXTRA})


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--networks", required=True,
                    help="folder of the .knet files")
    ap.add_argument("--out", required=True, help="output folder")
    a = ap.parse_args(argv)
    out = Path(a.out)
    out.mkdir(parents=True, exist_ok=True)
    metas = {}
    for name, (fn, th) in CONFIGS.items():
        text, meta = generate(Path(a.networks) / fn, th, name)
        (out / ("net_%s.h" % name)).write_text(text)
        (out / ("net_%s_meta.h" % name)).write_text(meta_header(meta, th))
        metas[name] = meta
    (out / "networks.json").write_text(json.dumps(metas, indent=1))


if __name__ == "__main__":
    main()
