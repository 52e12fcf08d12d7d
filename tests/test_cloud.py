"""The uneven batch's third network, cloud (452 species with the opaque
GRAIN0 and GRAIN-, 4429 reactions: networks/cloud.knet; opaque species
per docs/network-format.md:60-66 and kromex/knet.py:98-126), pinned
against the reference's own generator and solver (tests/golden/
cloud_*.npz, written by tests/golden/make_golden.py cloud):

- the generated tables: species, charges, element compositions and
  masses, the static-pivot sparse LU's tables (pivot order, rows of L + U
  with the fill, diagonals, flops) equal the reference generator's;
- the position table of the sparse Jacobian (LT_JP): every (term,
  scatter row) pair of the loop tables lands on the pattern entry of its
  own (row, column), decoded through the reference's tables;
- the wave form's gather tables against the loop form's scatter, bit for
  bit on random fluxes and terms; its rhs runs cover each contribution
  once, balanced over the lanes; its column tables of the factors against
  the reference's pattern;
- rates within 1e-9 relative (the same zeros), the rhs within 1e-10 of
  its row scale, the analytic Jacobian within 1e-12 of its scale;
- whole solves: the oracle against the reference's solve_cells (64 cells
  of the uneven batch, rtol 1e-6, linsol auto = the sparse LU), equal
  flags, 3e-4 (the uneven batch's bar, as primordialZ; 1e-4 above 1e-10
  of the cell total); the oracle's
  restatement bitwise against the reference's integrator on 8 cells;
- the same per-cell code the kernels run (kx::plain_cell), compiled for
  the CPU (tests/plain_host.cpp): the reference's steps exactly on the
  fixture's cheapest cells;
- GPU: the plain kernel's strided form (state, rhs and the sparse
  Jacobian in the interleaved workspace, the rates cached per solve)
  against the reference's solves and the oracle, and the lane queue
  bitwise equal to it, on the fixture's cheapest cells.
"""
import numpy as np
import pytest

import cells
import krome
from conftest import golden, oracle_opts, rel_err

NET = "cloud"
# the uneven batch's bar (rtol 1e-6, long integrations: up to 6843 steps),
# as primordialZ's: last-bit differences of the rate expressions move
# trace species (one cell's species at 1e-11 of the cell total by 2.7e-4);
# above 1e-10 of the cell total, 1e-4
TOL = 3e-4
TOL_MAJOR = 1e-4


# GPU tests take the fixture's cheapest cells: a few cells occupy one
# wave, whose lanes then run each step at the device memory's latency
# (the strided form reads and writes its ~1.5 MB of state per step in the
# workspace); the full batch is the bandwidth-bound case the benchmark
# measures
NGPU = 8


def _cheapest(d, m=NGPU):
    return np.sort(np.argsort(d["stats_bitwise"][:, 0])[:m])


def _opts(d):
    o = d["opts"]
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


def _generated_table(name):
    """a constexpr int table of the generated header"""
    from pathlib import Path
    text = (Path(__file__).resolve().parent.parent / "krome-x_amd" / "csrc"
            / "generated" / "net_cloud.h").read_text()
    key = "static constexpr int %s[" % name
    i = text.index(key)
    body = text[text.index("{", i) + 1:text.index("}", i)]
    return np.array([int(v) for v in body.split(",")])


def test_tables_equal_the_reference_generator():
    meta = cells.network_meta(NET)
    t = golden("cloud_tables.npz")
    assert list(t["element_names"]) == meta["elements"]
    assert np.array_equal(t["species_charge"], meta["charge"])
    assert np.allclose(t["species_mass_g"], meta["mass_g"], rtol=1e-12,
                       atol=0)
    assert np.array_equal(t["species_elements"], np.array(meta["natoms"]))
    # the opaque grains: no atoms, the grain anion one electron's mass
    sp = meta["species"]
    g0, gm = sp.index("GRAIN0"), sp.index("GRAIN-")
    assert not np.array(meta["natoms"])[[g0, gm]].any()
    assert meta["mass_g"][g0] == 0.0 and meta["mass_g"][gm] > 0.0
    s = golden("cloud_sparse.npz")
    for k in ("perm", "rp", "ci", "di"):
        assert np.array_equal(_generated_table("sp_" + k), s["sp_" + k]), k
    assert int(s["sparse_flops"]) == meta["sparse"]["flops"]


def test_sparse_jacobian_positions():
    """LT_JP against the reference's pattern: each (term, scatter row)
    pair of the loop tables, in loop order, names the position whose
    (row, column) the reference's tables decode to that pair's"""
    s = golden("cloud_sparse.npz")
    perm, rp, ci = s["sp_perm"], s["sp_rp"], s["sp_ci"]
    row_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    jp = _generated_table("LT_JP")
    jt_off, jt_col = _generated_table("LT_JT_OFF"), _generated_table(
        "LT_JT_COL")
    sc_off, sc_row = _generated_table("LT_SC_OFF"), _generated_table(
        "LT_SC_ROW")
    rows, cols = [], []
    for r in range(len(jt_off) - 1):
        for t in range(jt_off[r], jt_off[r + 1]):
            for q in range(sc_off[r], sc_off[r + 1]):
                rows.append(sc_row[q])
                cols.append(jt_col[t])
    assert len(jp) == len(rows)
    assert np.array_equal(perm[row_of[jp]], rows)
    assert np.array_equal(perm[ci[jp]], cols)


def _scatter_rows(vals, off, row, sgn, nrow):
    """the loop form's scatter: for each source in order, its rows +/- v"""
    f = [0.0] * nrow
    for r in range(len(off) - 1):
        v = float(vals[r])
        for q in range(off[r], off[r + 1]):
            f[row[q]] = f[row[q]] + v if sgn[q] > 0 else f[row[q]] - v
    return f


def _signed_sum(tab, e0, e1, vals, v=0.0):
    for e in range(e0, e1):
        x = tab[e]
        v = v + float(vals[x]) if x >= 0 else v - float(vals[~x])
    return v


def test_gather_tables_equal_the_scatter():
    """the wave form's gather tables (LT_ROW_*, LT_JPOS_*): each species
    row's and each pattern position's signed sum, in table order from 0,
    gives the loop form's scatter bit for bit on random fluxes and terms
    of mixed signs and magnitudes; the rhs runs (LT_SEG_*) cover every
    contribution once, are balanced over the 64 lanes, and give the
    scatter's bits on rows of one run and its value within roundoff on
    the longer rows"""
    rng = np.random.default_rng(7)
    sc_off, sc_row = _generated_table("LT_SC_OFF"), _generated_table(
        "LT_SC_ROW")
    sc_sgn = _generated_table("LT_SC_SGN")
    row_off, row_rs = _generated_table("LT_ROW_OFF"), _generated_table(
        "LT_ROW_RS")
    ns = len(row_off) - 1
    fl = rng.standard_normal(len(sc_off) - 1) * 10.0 ** rng.uniform(
        -30, 5, len(sc_off) - 1)
    want = _scatter_rows(fl, sc_off, sc_row, sc_sgn, ns)
    got = [_signed_sum(row_rs, row_off[i], row_off[i + 1], fl)
           for i in range(ns)]
    assert np.array_equal(np.array(got), np.array(want))
    # the runs
    lo, hi = _generated_table("LT_SEG_LO"), _generated_table("LT_SEG_HI")
    sid, rseg = _generated_table("LT_SEG_ID"), _generated_table(
        "LT_RSEG_OFF")
    assert len(lo) % 64 == 0
    cover = np.zeros(len(row_rs), dtype=int)
    part = {}
    for s_, a, b in zip(sid, lo, hi):
        if s_ < 0:
            assert a == b
            continue
        cover[a:b] += 1
        assert b - a <= 32
        part[int(s_)] = _signed_sum(row_rs, a, b, fl)
    assert (cover == 1).all() and sorted(part) == list(range(rseg[-1]))
    load = (hi - lo).reshape(-1, 64).sum(0)
    assert load.max() <= int(np.ceil(load.mean())) + 32
    for i in range(ns):
        runs = [part[q] for q in range(rseg[i], rseg[i + 1])]
        v = runs[0]
        for x in runs[1:]:
            v = v + x
        if len(runs) == 1:
            assert v == want[i]
        else:
            scale = sum(abs(float(fl[r if r >= 0 else ~r]))
                        for r in row_rs[row_off[i]:row_off[i + 1]])
            assert abs(v - want[i]) <= 1e-14 * scale
    # the Jacobian: terms scattered through LT_JP in loop order
    jt_off = _generated_table("LT_JT_OFF")
    jp = _generated_table("LT_JP")
    nt = jt_off[-1]
    dt = rng.standard_normal(nt) * 10.0 ** rng.uniform(-30, 5, nt)
    npos = int(jp.max()) + 1
    want = [0.0] * npos
    m = 0
    for r in range(len(jt_off) - 1):
        for t in range(jt_off[r], jt_off[r + 1]):
            for q in range(sc_off[r], sc_off[r + 1]):
                want[jp[m]] = (want[jp[m]] + float(dt[t]) if sc_sgn[q] > 0
                               else want[jp[m]] - float(dt[t]))
                m += 1
    pos_off, pos_ts = _generated_table("LT_JPOS_OFF"), _generated_table(
        "LT_JPOS_TS")
    got = [_signed_sum(pos_ts, pos_off[q], pos_off[q + 1], dt)
           for q in range(npos)]
    assert np.array_equal(np.array(got), np.array(want))


def test_factor_column_tables():
    """the wave form's column tables of the static-pivot factors (LT_LC*,
    LT_UC*) against the reference's row-ordered pattern: every strict
    lower and upper entry appears once, in its own column, rows increasing
    within a column, at its position; the per-position tables LT_PD and
    LT_PE are the pivot position and the U-row end of each entry's
    column"""
    sp = golden("cloud_sparse.npz")
    rp, ci, di = sp["sp_rp"], sp["sp_ci"], sp["sp_di"]
    ns = len(di)
    row_of = np.repeat(np.arange(ns), np.diff(rp))
    for tag, lower in (("LC", True), ("UC", False)):
        cp = _generated_table("LT_%sP" % tag)
        cr = _generated_table("LT_%sR" % tag)
        cq = _generated_table("LT_%sQ" % tag)
        part = np.where(ci < row_of)[0] if lower else \
            np.where(ci > row_of)[0]
        assert len(cp) == ns + 1 and cp[-1] == len(part) == len(cq)
        assert sorted(cq.tolist()) == part.tolist()
        for k in range(ns):
            q = cq[cp[k]:cp[k + 1]]
            assert (ci[q] == k).all()
            assert np.array_equal(row_of[q], cr[cp[k]:cp[k + 1]])
            assert (np.diff(cr[cp[k]:cp[k + 1]]) > 0).all()
    assert np.array_equal(_generated_table("LT_PD"), di[ci])
    assert np.array_equal(_generated_table("LT_PE"), rp[ci + 1])


def test_functions_vs_reference_generated_code(oracle):
    g = golden("cloud_funcs.npz")
    ns = g["n"].shape[1]
    k, f = oracle.funcs(NET, g["n"], g["T"], g["par"], 0)
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)
    nz = kr != 0.0
    assert np.max(np.abs(k[nz] - kr[nz]) / np.abs(kr[nz])) < 1e-9
    fr = g["f"][:, :ns]
    scale = np.abs(fr).max(axis=1, keepdims=True)
    assert np.max(np.abs(f[:, :ns] - fr) / scale) < 1e-10


def test_jacobian_vs_reference(oracle):
    import ctypes
    g = golden("cloud_funcs.npz")
    m = g["J"].shape[0]
    n, T, par = g["n"][:m], g["T"][:m], g["par"][:m]
    ns = n.shape[1]
    J = np.zeros((m, ns * ns))
    D = ctypes.POINTER(ctypes.c_double)
    oracle.lib.kxo_jac(21, m, np.ascontiguousarray(n).ctypes.data_as(D),
                       np.ascontiguousarray(T).ctypes.data_as(D),
                       np.ascontiguousarray(par).ctypes.data_as(D),
                       J.ctypes.data_as(D))
    J = J.reshape(m, ns, ns).transpose(0, 2, 1)
    Jr = g["J"]
    scale = np.abs(Jr).max(axis=(1, 2), keepdims=True)
    assert np.max(np.abs(J - Jr) / scale) < 1e-12


def test_oracle_vs_reference_solves(oracle):
    d = golden("cloud_solve.npz")
    r, n, T, fl, st = oracle.solve(NET, d["n0"], d["T0"], d["dt"], d["par"],
                                   _opts(d))
    assert np.array_equal(fl, d["flag_bitwise"])
    assert rel_err(n, d["n_bitwise"]) < TOL
    assert rel_err(n, d["n_bitwise"], 1e-10) < TOL_MAJOR
    s, sr = st[:, 0].sum(), d["stats_bitwise"][:, 0].sum()
    assert abs(s - sr) < 0.05 * sr


def test_oracle_bitwise_vs_reference_integrator(oracle, oracle_ref):
    d = golden("cloud_solve.npz")
    sl = slice(0, 8)
    o = _opts(d)
    a = oracle.solve(NET, d["n0"][sl], d["T0"][sl], d["dt"][sl],
                     d["par"][sl], o)
    b = oracle_ref.solve(NET, d["n0"][sl], d["T0"][sl], d["dt"][sl],
                         d["par"][sl], o)
    assert a[0] == b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def _host_lib():
    import subprocess
    from conftest import REPO
    out = REPO / "tests" / "_build" / "libplain_host.so"
    src = REPO / "tests" / "plain_host.cpp"
    hdr = REPO / "krome-x_amd" / "csrc"
    if not out.exists() or out.stat().st_mtime < max(
            f.stat().st_mtime for f in [src, *hdr.glob("*.h"),
                                        hdr / "generated" / "net_cloud.h"]):
        out.parent.mkdir(exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                        "-shared", "-fPIC", "-I", str(hdr), "-I",
                        str(hdr / "generated"), str(src), "-o", str(out)],
                       check=True)
    import ctypes
    return ctypes.CDLL(str(out))


def test_device_code_on_host_vs_reference():
    """the plain kernel's own per-cell code (kx::plain_cell, the strided
    form for cloud) compiled for the CPU: on the fixture's cheapest
    cells, the reference's steps exactly and its states to 1e-12"""
    import ctypes
    lib = _host_lib()
    d = golden("cloud_solve.npz")
    sel = _cheapest(d)
    n = np.ascontiguousarray(d["n0"][sel])
    T = np.ascontiguousarray(d["T0"][sel])
    o = np.array([d["opts"][1], d["opts"][2], d["opts"][3], d["opts"][6],
                  0, 0], dtype=np.float64)
    fl = np.zeros(len(sel), dtype=np.intc)
    st = np.zeros((len(sel), 9), dtype=np.int64)
    D = ctypes.POINTER(ctypes.c_double)
    r = lib.plain_host_solve(
        1, len(sel), n.ctypes.data_as(D), T.ctypes.data_as(D),
        np.ascontiguousarray(d["dt"][sel]).ctypes.data_as(D),
        np.ascontiguousarray(d["par"][sel]).ctypes.data_as(D),
        o.ctypes.data_as(D), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        st.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
    assert r == 0 and (fl == 0).all()
    assert np.array_equal(st[:, 0], d["stats_bitwise"][sel, 0])
    assert rel_err(n, d["n_bitwise"][sel]) < 1e-12


@pytest.mark.gpu
def test_gpu_plain_kernel_vs_reference(oracle):
    d = golden("cloud_solve.npz")
    sel = _cheapest(d)
    net = krome.load(NET)
    o = d["opts"]
    opts = net.options(evolve_T=0, rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=0, jac=0, max_steps=int(o[6]),
                       device_kernel="plain")
    print("cloud: plain kernel, cells %s" % sel.tolist(), flush=True)
    res = net.solve(d["n0"][sel], d["T0"][sel], d["dt"][sel],
                    par=d["par"][sel], opts=opts, check=False)
    assert res.ierr == 0
    assert np.array_equal(res.flags, d["flag_bitwise"][sel])
    assert rel_err(res.n, d["n_bitwise"][sel]) < TOL
    assert rel_err(res.n, d["n_bitwise"][sel], 1e-10) < TOL_MAJOR
    _, n, _, _, _ = oracle.solve(NET, d["n0"][sel], d["T0"][sel],
                                 d["dt"][sel], d["par"][sel], _opts(d))
    assert rel_err(res.n, n) < TOL


@pytest.mark.gpu
def test_gpu_refusals():
    """the strided form serves Tgas fixed with the analytic Jacobian and
    the sparse LU; the rest is refused before any work"""
    d = golden("cloud_solve.npz")
    net = krome.load(NET)
    for kw in ({"jac": 2}, {"linsol": 1}):
        opts = net.options(rtol=1e-6, **kw)
        res = net.solve(d["n0"][:4], d["T0"][:4], d["dt"][:4],
                        par=d["par"][:4], opts=opts, check=False)
        assert res.ierr == -104 and (res.flags == -104).all(), kw
        assert np.array_equal(res.n, d["n0"][:4])


@pytest.mark.gpu
def test_gpu_lane_queue_equals_plain():
    """device_kernel lanes: the lane queue (each lane solves cells with
    the plain kernel's code, taking the next from the device queue):
    bitwise the plain kernel's results"""
    d = golden("cloud_solve.npz")
    net = krome.load(NET)
    o = d["opts"]
    sel = _cheapest(d)
    res = {}
    for kern in ("plain", "lanes"):
        opts = net.options(evolve_T=0, rtol=o[1], atol_n=o[2], atol_T=o[3],
                           max_steps=int(o[6]), device_kernel=kern)
        print("cloud:", kern, flush=True)
        res[kern] = net.solve(d["n0"][sel], d["T0"][sel], d["dt"][sel],
                              par=d["par"][sel], opts=opts, check=False)
    assert res["lanes"].ierr == 0
    assert np.array_equal(res["lanes"].flags, res["plain"].flags)
    assert np.array_equal(res["lanes"].n, res["plain"].n)


@pytest.mark.gpu
@pytest.mark.parametrize("cpw", [1, 4, 16])
def test_gpu_wave_per_cell_vs_reference(oracle, cpw):
    """device_kernel persistent on cloud: a wave per cell (the rhs, the
    Jacobian, the sparse LU and its solves split over the 64 lanes, the
    norms by the butterfly), cpw cells per workgroup, every cell of the
    fixture: the reference's flags, its states within the bar; the same
    results for every workgroup shape"""
    d = golden("cloud_solve.npz")
    net = krome.load(NET)
    o = d["opts"]
    opts = net.options(evolve_T=0, rtol=o[1], atol_n=o[2], atol_T=o[3],
                       max_steps=int(o[6]), device_kernel="persistent",
                       cells_per_workgroup=cpw)
    print("cloud: a wave per cell, %d per workgroup" % cpw, flush=True)
    res = net.solve(d["n0"], d["T0"], d["dt"], par=d["par"], opts=opts,
                    check=False)
    assert res.ierr == 0
    assert np.array_equal(res.flags, d["flag_bitwise"])
    assert rel_err(res.n, d["n_bitwise"]) < TOL
    assert rel_err(res.n, d["n_bitwise"], 1e-10) < TOL_MAJOR
    _, n, _, _, _ = oracle.solve(NET, d["n0"], d["T0"], d["dt"], d["par"],
                                 _opts(d))
    assert rel_err(res.n, n) < TOL
    if cpw != 4:
        base = net.solve(d["n0"], d["T0"], d["dt"], par=d["par"],
                         opts=net.options(evolve_T=0, rtol=o[1],
                                          atol_n=o[2], atol_T=o[3],
                                          max_steps=int(o[6]),
                                          device_kernel="persistent"),
                         check=False)
        assert np.array_equal(base.n, res.n)
