"""ctypes wrapper of the CPU oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It loads oracle/_build/liboracle.so (parity build, no FMA contraction) or
liboracle_fast.so (timed baseline build) produced by oracle/Makefile; the
environment variable SPHGPU_ORACLE_LIB names another parity build to load
instead (the sanitized one).

Array conventions follow Spheral's Field storage: one row per node, Vector =
D doubles, SymTensor = D(D+1)/2 doubles (xx,xy,xz,yy,yz,zz), Tensor = D*D doubles
(row major).  Node order is the reference's: internal nodes first, then ghosts.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [("Cl", C.c_double), ("Cq", C.c_double), ("eps2", C.c_double),
                ("epsTensile", C.c_double),
                ("nPerh", C.c_double), ("hmin", C.c_double), ("hmax", C.c_double),
                ("hminratio", C.c_double),
                ("XSPH", C.c_int), ("correctVelocityGradient", C.c_int),
                ("evolveTotalEnergy", C.c_int),
                ("linearInExpansion", C.c_int), ("quadraticInExpansion", C.c_int)]


class State(C.Structure):
    _fields_ = [(n, _dp) for n in ("pos", "vel", "mass", "rho", "eps", "P", "cs", "H", "omega",
                                   "fCl", "fCq", "fshear")] + [("list", _i32p)]


OUTPUT_FIELDS = ["rhoSum", "normalization", "DxDt", "DrhoDt", "DvDt", "DepsDt", "DvDx",
                 "localDvDx", "M", "localM", "DHDt", "Hideal", "maxViscousPressure",
                 "effViscousPressure", "viscousWork", "XSPHWeightSum", "XSPHDeltaV",
                 "weightedNeighborSum", "massSecondMoment"]


class Plane(C.Structure):
    _fields_ = [("point", C.c_double * 3), ("normal", C.c_double * 3)]


class Boundary(C.Structure):
    """kind 0: ReflectingBoundary(p1); kind 1: PeriodicBoundary(p1, p2)."""
    _fields_ = [("kind", C.c_int), ("p1", Plane), ("p2", Plane)]


def plane(point, normal):
    pt = list(point) + [0.0] * (3 - len(point))
    nm = list(normal) + [0.0] * (3 - len(normal))
    return Plane((C.c_double * 3)(*pt), (C.c_double * 3)(*nm))


def reflecting(point, normal):
    return Boundary(0, plane(point, normal), plane(point, normal))


def periodic(point1, normal1, point2, normal2):
    return Boundary(1, plane(point1, normal1), plane(point2, normal2))


def box_boundaries(dim, lo=0.0, hi=1.0, kind="reflecting"):
    """Planes at lo and hi on every axis, inward normals, in x, y, z order."""
    out = []
    for a in range(dim):
        e = [0.0] * dim
        e[a] = 1.0
        p_lo = [0.0] * dim
        p_lo[a] = lo
        p_hi = [0.0] * dim
        p_hi[a] = hi
        if kind == "reflecting":
            out.append(reflecting(p_lo, e))
            out.append(reflecting(p_hi, [-x for x in e]))
        else:
            out.append(periodic(p_lo, e, p_hi, [-x for x in e]))
    return out


class Outputs(C.Structure):
    _fields_ = [(n, _dp) for n in OUTPUT_FIELDS] + [("ncalc", _i64p), ("pairAcc", _dp)]


def ncomp(field, dim):
    nsym = dim * (dim + 1) // 2
    return {"DxDt": dim, "DvDt": dim, "XSPHDeltaV": dim,
            "DvDx": dim * dim, "localDvDx": dim * dim, "M": dim * dim, "localM": dim * dim,
            "DHDt": nsym, "Hideal": nsym, "massSecondMoment": nsym}.get(field, 1)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(BUILD, name)
        # SPHGPU_ORACLE_LIB: another build of the parity library (the sanitized one)
        if not fast and os.environ.get("SPHGPU_ORACLE_LIB"):
            path = os.environ["SPHGPU_ORACLE_LIB"]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_table_create.restype = C.c_void_p
        L.orc_table_create.argtypes = [C.c_int, C.c_int]
        L.orc_table_destroy.argtypes = [C.c_void_p]
        L.orc_table_npoints.argtypes = [C.c_void_p]
        L.orc_table_export.argtypes = [C.c_void_p, _dp]
        for f in ("orc_table_kernel", "orc_table_grad", "orc_table_grad2"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_table_nperh_from_wsum.restype = C.c_double
        L.orc_table_nperh_from_wsum.argtypes = [C.c_void_p, C.c_double]
        L.orc_bspline.restype = C.c_double
        L.orc_bspline.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_connectivity.restype = C.c_int64
        L.orc_connectivity.argtypes = [C.c_int, C.c_int64, C.c_int64, _dp, _dp, C.c_double, C.c_int,
                                       _i64p, _i32p]
        L.orc_sum_density.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, _dp, _dp, _dp,
                                      _i64p, _i32p, _dp]
        L.orc_eval_derivs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(Params), C.c_int64,
                                      C.c_int64, C.POINTER(State), _i64p, _i32p, C.POINTER(Outputs)]
        L.orc_compatible_depsdt.argtypes = [C.c_int, C.c_int64, C.c_int64, _i64p, _i32p, _dp, _dp, _dp, _dp, _dp,
                                            C.c_double, _dp]
        L.orc_omega_gradh.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, _dp, _dp, _i64p, _i32p, _dp]
        L.orc_ideal_h_pass.argtypes = [C.c_void_p, C.c_int, C.POINTER(Params), C.c_int64, _dp, _dp, _dp, _i64p, _i32p,
                                       C.c_double, _i32p, _dp, _dp]
        L.orc_ideal_h_pass.restype = C.c_double
        L.orc_gamma_law.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                    _dp, _dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_sum_density_lists.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, _dp, _dp, _dp,
                                            _i64p, _i32p, _i32p, C.c_int, _dp]
        L.orc_set_ghosts_lists.restype = C.c_int64
        L.orc_set_ghosts_lists.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int, C.POINTER(Boundary), C.c_double,
                                           C.c_int64, _i32p, _i32p, _i32p, _i32p]
        L.orc_set_ghosts.restype = C.c_int64
        L.orc_set_ghosts.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int, C.POINTER(Boundary), C.c_double,
                                     C.c_int64, _i32p, _i32p, _i32p]
        L.orc_apply_ghosts.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int, C.POINTER(Boundary), _i32p, _i32p,
                                       C.c_int, _dp]
        L.orc_max_eigen_n.argtypes = [C.c_int, C.c_int64, _dp, _dp]
        L.orc_enforce_boundaries.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(Boundary), _dp, _dp, _dp]
        L.orc_min_eigen.restype = C.c_double
        L.orc_min_eigen.argtypes = [C.c_int, _dp]
        L.orc_cull_map.restype = C.c_int64
        L.orc_cull_map.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _i32p, _i32p]
        _libs[name] = L
    return _libs[name]


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct) if a is not None else C.cast(None, ct)


class Table:
    """TableKernel(BSplineKernel<dim>, npoints) restated (TableKernel.cc:211-305)."""

    def __init__(self, dim, npoints=1000, fast=False):
        self.L = lib(fast)
        self.dim = dim
        self.h = self.L.orc_table_create(dim, npoints)
        if not self.h:
            raise ValueError("bad table arguments")
        self.n = npoints

    def __del__(self):
        try:
            self.L.orc_table_destroy(self.h)
        except Exception:
            pass

    def arrays(self):
        out = np.zeros(11 * self.n)
        self.L.orc_table_export(self.h, _p(out))
        names = ["Ak", "Bk", "Ck", "Ag", "Bg", "Cg", "Ag2", "Bg2", "Cg2", "nperh", "wsum"]
        return {k: out[i * self.n:(i + 1) * self.n].copy() for i, k in enumerate(names)}

    def W(self, eta, Hdet=1.0):
        return self.L.orc_table_kernel(self.h, eta, Hdet)

    def gradW(self, eta, Hdet=1.0):
        return self.L.orc_table_grad(self.h, eta, Hdet)

    def grad2W(self, eta, Hdet=1.0):
        return self.L.orc_table_grad2(self.h, eta, Hdet)

    def nperh_from_wsum(self, w):
        return self.L.orc_table_nperh_from_wsum(self.h, w)


def bspline(dim, which, eta, Hdet=1.0):
    return lib().orc_bspline(dim, which, eta, Hdet)


def connectivity(dim, n_internal, pos, H, ext=2.0, brute=False, fast=False):
    """Exact ConnectivityMap rows (CSR, ascending neighbour index)."""
    L = lib(fast)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    H = np.ascontiguousarray(H, dtype=np.float64)
    n_total = pos.shape[0]
    off = np.zeros(n_internal + 1, dtype=np.int64)
    total = L.orc_connectivity(dim, n_internal, n_total, _p(pos), _p(H), ext, int(brute), _p(off, _i64p),
                               C.cast(None, _i32p))
    nbr = np.zeros(max(total, 1), dtype=np.int32)
    L.orc_connectivity(dim, n_internal, n_total, _p(pos), _p(H), ext, int(brute), _p(off, _i64p),
                       _p(nbr, _i32p))
    return off, nbr[:total]


def list_order(n_internal, lists, off, nbr, receive):
    """Rows of several NodeLists in the reference's order.  lists: NodeList of
    every node (internal rows of each NodeList contiguous and in NodeList
    order, ghosts after all internal rows).  receive=False: ConnectivityMap's
    order, by (NodeList, index within the NodeList) (the scatter walks it);
    receive=True: the order in which the scatter delivers node i's terms, the
    internal neighbours that own their pair (lower index) first, then i's own
    pairs in ConnectivityMap order (the gather walks it)."""
    lists = np.asarray(lists, dtype=np.int64)
    n_tot = lists.shape[0]
    idx = np.arange(n_tot)
    # rank in (NodeList, index within the NodeList) order: a NodeList's ghosts
    # follow its internal rows
    ghost = idx >= n_internal
    rank = np.lexsort((idx, ghost, lists))
    pos_of = np.empty(n_tot, dtype=np.int64)
    pos_of[rank] = idx
    out = nbr.copy()
    for i in range(n_internal):
        row = nbr[off[i]:off[i + 1]].astype(np.int64)
        key = pos_of[row]
        if receive:
            owned_by_j = (row < n_internal) & (row < i)
            key = np.where(owned_by_j, row, n_tot + key)
        out[off[i]:off[i + 1]] = row[np.argsort(key, kind="stable")]
    return out


def sum_density(table, n_internal, pos, mass, H, off, nbr, gather=False, lists=None, sum_all=True):
    """computeSPHSumMassDensity restated.  lists (NodeList of every node) and
    sum_all (sumMassDensityOverAllNodeLists) for several NodeLists; the rows
    are reordered as the reference walks them."""
    L = table.L
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rho = np.zeros(n_internal)
    args = (_p(np.ascontiguousarray(mass, dtype=np.float64)), _p(np.ascontiguousarray(H, dtype=np.float64)))
    if lists is None:
        L.orc_sum_density(table.h, table.dim, int(gather), n_internal, pos.shape[0], _p(pos), *args,
                          _p(off, _i64p), _p(nbr, _i32p), _p(rho))
        return rho
    la = np.ascontiguousarray(lists, dtype=np.int32)
    rows = np.ascontiguousarray(list_order(n_internal, la, off, nbr, receive=gather), dtype=np.int32)
    L.orc_sum_density_lists(table.h, table.dim, int(gather), n_internal, pos.shape[0], _p(pos), *args,
                            _p(off, _i64p), _p(rows, _i32p), _p(la, _i32p), int(bool(sum_all)), _p(rho))
    return rho


def omega_gradh(table, n_internal, pos, H, off, nbr, lists=None):
    """computeSPHOmegaGradhCorrection restated (internal nodes): the owner
    loop over every pair once, all NodeLists alike
    (src/SPH/computeSPHOmegaGradhCorrection.cc:55-130).  lists (NodeList of
    every node): several NodeLists; node i then walks its own pairs in
    ConnectivityMap order, by (NodeList, index within the NodeList).  With
    each NodeList's internal rows contiguous in NodeList order and the ghosts
    after all of them, the owner of a pair (ConnectivityMapInline.hh:182-184)
    is its end of lower row index, as with one NodeList."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    om = np.zeros(n_internal)
    if lists is not None:
        nbr = np.ascontiguousarray(list_order(n_internal, lists, off, nbr, receive=False), dtype=np.int32)
    table.L.orc_omega_gradh(table.h, table.dim, n_internal, pos.shape[0], _p(pos),
                            _p(np.ascontiguousarray(H, dtype=np.float64)), _p(off, _i64p), _p(nbr, _i32p), _p(om))
    return om


def make_params(**kw):
    d = dict(Cl=1.0, Cq=1.0, eps2=1.0e-2, epsTensile=0.0, nPerh=1.51, hmin=1.0e-20, hmax=1.0e20,
             hminratio=0.1, XSPH=1, correctVelocityGradient=1, evolveTotalEnergy=0,
             linearInExpansion=0, quadraticInExpansion=0)
    d.update(kw)
    return Params(**d)


def eval_derivs(table, params, n_internal, state, off, nbr, gather=False, pair_acc=False, lists=None):
    """SPHHydroBase::evaluateDerivatives restated.  state: dict of numpy arrays
    (pos, vel, mass, rho, eps, P, cs, H, omega, optional fCl, fCq, fshear) over all
    nodes.  Returns dict of per-internal-node outputs (+ 'ncalc' for scatter).
    lists: NodeList of every node (several NodeLists); the rows are reordered
    as the reference walks them."""
    dim = table.dim
    keep = {}
    st = State()
    for name, _ in State._fields_:
        if name == "list":
            continue
        a = state.get(name)
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[name] = a
        setattr(st, name, _p(a))
    if lists is not None:
        keep["list"] = np.ascontiguousarray(lists, dtype=np.int32)
        st.list = _p(keep["list"], _i32p)
        nbr_in = nbr
        nbr = np.ascontiguousarray(list_order(n_internal, keep["list"], off, nbr, receive=gather), dtype=np.int32)
    else:
        st.list = C.cast(None, _i32p)
    outs = {f: np.zeros((n_internal, ncomp(f, dim))) for f in OUTPUT_FIELDS}
    ncalc = np.zeros(1, dtype=np.int64)
    pa = np.zeros((max(int(off[-1]), 1), dim)) if pair_acc else None
    o = Outputs(**{f: _p(outs[f]) for f in OUTPUT_FIELDS}, ncalc=_p(ncalc, _i64p), pairAcc=_p(pa))
    n_total = keep["pos"].shape[0]
    table.L.orc_eval_derivs(table.h, dim, int(gather), C.byref(params), n_internal, n_total, C.byref(st),
                            _p(off, _i64p), _p(nbr, _i32p), C.byref(o))
    res = {f: (v[:, 0] if v.shape[1] == 1 else v) for f, v in outs.items()}
    if not gather:
        res["ncalc"] = int(ncalc[0])
    if pair_acc:
        # each node's pair accelerations in the order they reach it (the
        # scatter's pushes); with several NodeLists that is the delivery order,
        # returned as rows in which they sit entry for entry
        res["pairAcc"] = pa[:int(off[-1])]
        if lists is not None:
            res["nbr_order"] = np.ascontiguousarray(list_order(n_internal, keep["list"], off, nbr_in, receive=True),
                                                    dtype=np.int32)
    return res


def ideal_h_pass(table, params, n_internal, pos, H, deltaH, off, nbr, tol, lists=None, mass=None, rho=None):
    """One iterateIdealH pass (SPH smoothing scale); H (all rows, internal ones
    replaced) and deltaH (internal rows) are updated in place.  Returns the
    largest change of the nodes computed.  lists, mass, rho (every node):
    several NodeLists; a neighbour of another NodeList then counts with
    m_j rho_i / (m_i rho_j) (iterateIdealH.cc:183-194) and node i walks its
    full connectivity in ConnectivityMap order."""
    assert H.flags.c_contiguous and deltaH.flags.c_contiguous and H.dtype == np.float64
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    la = ma = ra = None
    if lists is not None:
        la = np.ascontiguousarray(lists, dtype=np.int32)
        ma = np.ascontiguousarray(mass, dtype=np.float64)
        ra = np.ascontiguousarray(rho, dtype=np.float64)
        nbr = np.ascontiguousarray(list_order(n_internal, la, off, nbr, receive=False), dtype=np.int32)
    return table.L.orc_ideal_h_pass(table.h, table.dim, C.byref(params), n_internal, _p(pos), _p(H), _p(deltaH),
                                    _p(off, _i64p), _p(nbr, _i32p), tol,
                                    _p(la, _i32p), _p(ma), _p(ra))


def compatible_depsdt(dim, n_internal, off, nbr, mass, vel, acc, DepsDt0, pair_acc, multiplier):
    """SpecificThermalEnergyPolicy restated: DepsDt of the internal nodes."""
    out = np.zeros(n_internal)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    mass, vel, acc, DepsDt0, pair_acc = f(mass), f(vel), f(acc), f(DepsDt0), f(pair_acc)
    lib().orc_compatible_depsdt(dim, n_internal, mass.shape[0], _p(off, _i64p), _p(nbr, _i32p), _p(mass), _p(vel),
                                _p(acc), _p(DepsDt0), _p(pair_acc), multiplier, _p(out))
    return out


def gamma_law(rho, eps, gamma, Pmin=-1e200, Pmax=1e200, floor=False):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    P = np.zeros_like(rho)
    cs = np.zeros_like(rho)
    lib().orc_gamma_law(rho.size, gamma, Pmin, Pmax, int(floor), _p(rho), _p(eps), _p(P), _p(cs))
    return P, cs


def num_threads(fast=True):
    return lib(fast).orc_num_threads()


def _barr(boundaries):
    arr = (Boundary * max(1, len(boundaries)))(*boundaries)
    return arr


def n_sub(boundaries):
    return sum(1 if b.kind == 0 else 2 for b in boundaries)


def set_ghosts(dim, pos, H, boundaries, ext=2.0, lists=None):
    """PlanarBoundary::setGhostNodes for each boundary in order over internal
    rows pos/H.  Returns (pos_all, H_all, control, sub_of, sub_start), and with
    lists (NodeList of every internal row: each NodeList selects its own ghosts)
    also the NodeList of every row."""
    L = lib()
    nsym = dim * (dim + 1) // 2
    n = pos.shape[0]
    cap = max(64, n)
    barr = _barr(boundaries)
    while True:
        P = np.zeros((n + cap, dim))
        Hh = np.zeros((n + cap, nsym))
        P[:n] = pos
        Hh[:n] = H
        control = np.zeros(cap, dtype=np.int32)
        sub_of = np.zeros(cap, dtype=np.int32)
        sub_start = np.zeros(n_sub(boundaries) + 1, dtype=np.int32)
        if lists is None:
            ng = L.orc_set_ghosts(dim, n, _p(P), _p(Hh), len(boundaries), barr, ext, cap, _p(control, _i32p),
                                  _p(sub_of, _i32p), _p(sub_start, _i32p))
        else:
            la = np.zeros(n + cap, dtype=np.int32)
            la[:n] = lists
            ng = L.orc_set_ghosts_lists(dim, n, _p(P), _p(Hh), len(boundaries), barr, ext, cap, _p(control, _i32p),
                                        _p(sub_of, _i32p), _p(sub_start, _i32p), _p(la, _i32p))
        if ng >= 0:
            if lists is None:
                return P[:n + ng], Hh[:n + ng], control[:ng], sub_of[:ng], sub_start
            return P[:n + ng], Hh[:n + ng], control[:ng], sub_of[:ng], sub_start, la[:n + ng]
        cap *= 2


FIELD_KIND = {"pos": 3, "vel": 1, "H": 2}


def max_eigen_n(dim, H):
    """SymTensor::eigenValues().maxElement() of every row."""
    a = np.ascontiguousarray(H, dtype=np.float64)
    out = np.zeros(a.shape[0])
    lib().orc_max_eigen_n(dim, a.shape[0], _p(a), _p(out))
    return out


def enforce_boundaries(dim, n_internal, boundaries, pos, vel, H):
    """Integrator::enforceBoundaries on the internal rows (in place)."""
    lib().orc_enforce_boundaries(dim, n_internal, len(boundaries), _barr(boundaries), _p(pos), _p(vel), _p(H))


def apply_ghosts(dim, n_internal, boundaries, control, sub_of, field, kind):
    """Boundary::applyGhostBoundary over every ghost row of `field` (in place).
    kind 0 scalar, 1 Vector, 2 SymTensor, 3 positions (updateGhostNodes)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    lib().orc_apply_ghosts(dim, n_internal, f.shape[0], len(boundaries), _barr(boundaries),
                           _p(np.ascontiguousarray(control, dtype=np.int32), _i32p),
                           _p(np.ascontiguousarray(sub_of, dtype=np.int32), _i32p), kind, _p(f))
    return f


def min_eigen(dim, h):
    """SymTensor::eigenValues().minElement() of a packed SymTensor."""
    a = np.ascontiguousarray(h, dtype=np.float64)
    return lib().orc_min_eigen(dim, _p(a))


def cull_map(n_internal, n_total, off, nbr, control):
    old2new = np.zeros(n_total, dtype=np.int32)
    m = lib().orc_cull_map(n_internal, n_total, _p(off, _i64p), _p(nbr, _i32p),
                           _p(np.ascontiguousarray(control, dtype=np.int32), _i32p), _p(old2new, _i32p))
    return int(m), old2new


STATE_KINDS = {"vel": 1, "mass": 0, "rho": 0, "eps": 0, "P": 0, "cs": 0, "omega": 0,
               "fCl": 0, "fCq": 0, "fshear": 0}


def ghost_pipeline(dim, state, boundaries, ext=2.0, cull=True, fast=False, lists=None):
    """Integrator::setGhostNodes + ConnectivityMap + culling + applyGhostBoundaries
    on internal-node state.  Returns (state_all, off, nbr, info) with ghosts
    appended after the internal rows and rows in the culled numbering.  With
    lists (NodeList of every internal row), each NodeList makes its own ghosts
    and state_all["list"] holds the NodeList of every row."""
    n = state["pos"].shape[0]
    if lists is None:
        pos, H, control, sub_of, sub_start = set_ghosts(dim, state["pos"], state["H"], boundaries, ext)
    else:
        pos, H, control, sub_of, sub_start, la = set_ghosts(dim, state["pos"], state["H"], boundaries, ext, lists)
    off, nbr = connectivity(dim, n, pos, H, ext, fast=fast)
    out = {"pos": pos, "H": H}
    if lists is not None:
        out["list"] = la
    for k, kind in STATE_KINDS.items():
        if state.get(k) is None:
            continue
        a = state[k]
        full = np.zeros((pos.shape[0],) + a.shape[1:])
        full[:n] = a
        out[k] = apply_ghosts(dim, n, boundaries, control, sub_of, full, kind)
    info = {"n_ghost_generated": pos.shape[0] - n, "control": control, "sub_of": sub_of,
            "sub_start": sub_start}
    if cull:
        m, old2new = cull_map(n, pos.shape[0], off, nbr, control)
        keep = old2new >= 0
        for k in list(out):
            out[k] = out[k][keep]
        gk = keep[n:]
        info["control"] = old2new[control[gk]]
        info["sub_of"] = sub_of[gk]
        info["old2new"] = old2new
        nbr = old2new[nbr]
    return out, off, nbr, info
