"""Python binding of libsphgpu.so, the MI355X SPH derivative pass.

This is a thin ctypes layer over the C-ABI in include/sphgpu.h; every compute
call runs the HIP kernels of the library.  There is no CPU fallback: if the
library is missing, importing works but creating a Sim raises.

Field names follow src/Hydro/HydroFieldNames.cc:11-63 of the reference.
"""
import ctypes as C
import os

import numpy as np

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libsphgpu.so")

# state fields (SPHGPU_F_*)
STATE_FIELDS = {
    "position": 0, "velocity": 1, "mass": 2, "mass density": 3, "specific thermal energy": 4,
    "H": 5, "pressure": 6, "sound speed": 7, "grad h corrections": 8,
    "Cl multiplier": 9, "Cq multiplier": 10, "Balsara shear correction": 11,
}
# derivative fields (SPHGPU_D_*), in the order evaluateDerivatives fetches them
DERIV_FIELDS = [
    "new mass density", "normalization", "delta position", "delta mass density",
    "delta velocity hydro", "delta specific thermal energy", "velocity gradient",
    "internal velocity gradient", "M SPH gradient correction", "local M SPH gradient correction",
    "delta H", "new H", "max viscous pressure", "effective viscous pressure", "viscous work rate",
    "XSPH weight sum", "XSPH delta vi", "weighted neighbor sum", "mass second moment",
]
# short aliases used by the tests / oracle
DERIV_ALIASES = [
    "rhoSum", "normalization", "DxDt", "DrhoDt", "DvDt", "DepsDt", "DvDx", "localDvDx", "M", "localM",
    "DHDt", "Hideal", "maxViscousPressure", "effViscousPressure", "viscousWork", "XSPHWeightSum",
    "XSPHDeltaV", "weightedNeighborSum", "massSecondMoment",
]

PRECISION_FP64 = 0
PRECISION_MIXED = 1
ROW_ORDER_REFERENCE = 0   # rows ascending in node index (reference summation order)
ROW_ORDER_SEARCH = 1      # deterministic device search order (no per-row sort)

ERRORS = {-1: "ARG", -2: "STATE", -3: "HIP", -4: "UNSUPPORTED", -5: "CAPACITY"}


class SphGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sphgpu error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class HydroParams(C.Structure):
    _fields_ = [("Cl", C.c_double), ("Cq", C.c_double), ("eps2", C.c_double), ("epsTensile", C.c_double),
                ("nPerh", C.c_double), ("hmin", C.c_double), ("hmax", C.c_double), ("hminratio", C.c_double),
                ("XSPH", C.c_int), ("correctVelocityGradient", C.c_int), ("evolveTotalEnergy", C.c_int),
                ("linearInExpansion", C.c_int), ("quadraticInExpansion", C.c_int),
                ("compatibleEnergyEvolution", C.c_int), ("precision", C.c_int)]


BOUNDARY_REFLECTING = 0
BOUNDARY_PERIODIC = 1


class Member(C.Structure):
    """sphgpu_member: one simulation of an ensemble."""
    _fields_ = [("n", C.c_int64), ("pos", _dp), ("vel", _dp), ("mass", _dp), ("rho", _dp), ("eps", _dp),
                ("H", _dp), ("gamma", C.c_double), ("tEnd", C.c_double)]


class Request(C.Structure):
    """sphgpu_request: a simulation arriving at a serving loop."""
    _fields_ = [("member", Member), ("arrival", C.c_double), ("outPos", _dp), ("outVel", _dp), ("outRho", _dp),
                ("outEps", _dp), ("outH", _dp), ("admitted", C.c_double), ("finished", C.c_double),
                ("steps", C.c_int64), ("status", C.c_int32)]


class ServeParams(C.Structure):
    _fields_ = [("hbmBudget", C.c_int64), ("maxMembers", C.c_int), ("ghostFactor", C.c_double),
                ("dtThreshold", C.c_double)]


# reasons an ensemble member fails (bits, SPHGPU_MEMBER_*)
MEMBER_OK, MEMBER_NONFINITE, MEMBER_BAD_H, MEMBER_DT_COLLAPSE = 0, 1, 2, 4


class Plane(C.Structure):
    _fields_ = [("point", C.c_double * 3), ("normal", C.c_double * 3)]


class Boundary(C.Structure):
    _fields_ = [("kind", C.c_int), ("plane1", Plane), ("plane2", Plane)]


def plane(point, normal):
    pt = list(point) + [0.0] * (3 - len(point))
    nm = list(normal) + [0.0] * (3 - len(normal))
    return Plane((C.c_double * 3)(*pt), (C.c_double * 3)(*nm))


def reflecting(point, normal):
    """ReflectingBoundary(GeomPlane(point, normal)); normal points inwards."""
    return Boundary(BOUNDARY_REFLECTING, plane(point, normal), plane(point, normal))


def periodic(point1, normal1, point2, normal2):
    """PeriodicBoundary(GeomPlane(point1, normal1), GeomPlane(point2, normal2))."""
    return Boundary(BOUNDARY_PERIODIC, plane(point1, normal1), plane(point2, normal2))


def box_boundaries(dim, lo=0.0, hi=1.0, kind="reflecting"):
    """Planes at lo and hi on every axis with inward normals, x then y then z
    (reflecting: lo plane then hi plane per axis; periodic: one pair per axis)."""
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


class IntegratorParams(C.Structure):
    _fields_ = [("cfl", C.c_double), ("dtMin", C.c_double), ("dtMax", C.c_double), ("dtGrowth", C.c_double),
                ("gamma", C.c_double), ("Pmin", C.c_double), ("Pmax", C.c_double), ("pressureFloor", C.c_int),
                ("cullGhostNodes", C.c_int), ("rhoMin", C.c_double), ("rhoMax", C.c_double),
                ("kernelExtent", C.c_double), ("useVelocityMagnitudeForDt", C.c_int),
                ("gradhCorrection", C.c_int)]


FLT_MAX = 3.4028234663852886e38


def integrator_params(**kw):
    """SynchronousRK2 / Integrator, GenericHydro, GammaLawGas and FluidNodeList
    defaults of the reference (see include/sphgpu.h)."""
    d = dict(cfl=0.25, dtMin=0.0, dtMax=FLT_MAX, dtGrowth=2.0, gamma=5.0 / 3.0, Pmin=-1.0e200, Pmax=1.0e200,
             pressureFloor=1, cullGhostNodes=1, rhoMin=1.0e-10, rhoMax=1.0e10, kernelExtent=2.0,
             useVelocityMagnitudeForDt=0, gradhCorrection=0)
    d.update(kw)
    return IntegratorParams(**d)


def hydro_params(**kw):
    """Defaults of the reference's SPH() factory for the B-spline kernel
    (src/SPH/SPHHydros.py:317-384) with compatible energy and grad-h off."""
    d = dict(Cl=1.0, Cq=1.0, eps2=1.0e-2, epsTensile=0.0, nPerh=1.51, hmin=1.0e-20, hmax=1.0e20,
             hminratio=0.1, XSPH=1, correctVelocityGradient=1, evolveTotalEnergy=0,
             linearInExpansion=0, quadraticInExpansion=0, compatibleEnergyEvolution=0,
             precision=PRECISION_FP64)
    d.update(kw)
    return HydroParams(**d)


_lib = None


def library_path():
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsphgpu.so not built ({LIB_PATH}); run llnl-spheral_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "sphgpu_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
            "sphgpu_destroy": (None, [vp]),
            "sphgpu_last_error": (C.c_char_p, [vp]),
            "sphgpu_set_table": (C.c_int, [vp, C.c_int, C.c_double, _dp]),
            "sphgpu_set_table_bspline": (C.c_int, [vp, C.c_int]),
            "sphgpu_get_table": (C.c_int, [vp, C.c_int, _dp]),
            "sphgpu_table_values": (C.c_int, [vp, C.c_double, C.c_double, _dp, _dp]),
            "sphgpu_set_nodes": (C.c_int, [vp, C.c_int64, C.c_int64]),
            "sphgpu_upload": (C.c_int, [vp, C.c_int, _dp, C.c_int64]),
            "sphgpu_download": (C.c_int, [vp, C.c_int, _dp, C.c_int64]),
            "sphgpu_set_row_order": (C.c_int, [vp, C.c_int]),
            "sphgpu_build_neighbors": (C.c_int, [vp, C.c_double]),
            "sphgpu_neighbor_count": (C.c_int, [vp, _i64p]),
            "sphgpu_download_neighbors": (C.c_int, [vp, _i64p, _i32p]),
            "sphgpu_sum_density": (C.c_int, [vp]),
            "sphgpu_set_precision": (C.c_int, [vp, C.c_int]),
            "sphgpu_evaluate_derivatives": (C.c_int, [vp, C.POINTER(HydroParams)]),
            "sphgpu_download_derivative": (C.c_int, [vp, C.c_int, _dp, C.c_int64]),
            "sphgpu_pair_count": (C.c_int, [vp, _i64p]),
            "sphgpu_last_timings": (C.c_int, [vp, _dp]),
            "sphgpu_synchronize": (C.c_int, [vp]),
            "sphgpu_copy_probe": (C.c_int, [vp, C.c_int64, C.c_int, _dp]),
            "sphgpu_gather_probe": (C.c_int, [vp, C.c_int, C.c_int, _dp]),
            "sphgpu_set_boundaries": (C.c_int, [vp, C.c_int, C.POINTER(Boundary)]),
            "sphgpu_eos_gamma_law": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_int]),
            "sphgpu_omega_gradh": (C.c_int, [vp]),
            "sphgpu_hydro_dt": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, _dp]),
            "sphgpu_rk2_step": (C.c_int, [vp, C.POINTER(HydroParams), C.POINTER(IntegratorParams), C.c_double, _dp]),
            "sphgpu_get_time": (C.c_int, [vp, _dp, _dp, _i64p]),
            "sphgpu_set_time": (C.c_int, [vp, C.c_double, C.c_double, C.c_int64]),
            "sphgpu_set_ghost_nodes": (C.c_int, [vp, C.c_double, C.c_int]),
            "sphgpu_apply_ghost_boundaries": (C.c_int, [vp]),
            "sphgpu_node_counts": (C.c_int, [vp, _i64p, _i64p]),
            "sphgpu_download_ghost_controls": (C.c_int, [vp, _i32p, _i32p]),
            "sphgpu_iterate_ideal_h": (C.c_int, [vp, C.POINTER(HydroParams), C.c_int, C.c_double, C.c_double, C.c_int,
                                                 C.POINTER(C.c_int), _dp]),
            "sphgpu_conservation": (C.c_int, [vp, _dp]),
            "sphgpu_get_dt_inputs": (C.c_int, [vp, C.c_int64, C.c_int64, _dp]),
            "sphgpu_set_dt_inputs": (C.c_int, [vp, C.c_int64, C.c_int64, _dp]),
            "sphgpu_ensemble_conservation": (C.c_int, [vp, C.c_int64, _dp]),
            "sphgpu_conservation_noalloc": (C.c_int, [vp, C.c_int64, C.c_int64, C.c_void_p]),
            "sphgpu_set_graphs": (C.c_int, [vp, C.c_int]),
            "sphgpu_set_node_lists": (C.c_int, [vp, C.c_int, _i64p, _dp, C.c_int]),
            "sphgpu_autotune": (C.c_int, [C.c_int]),
            "sphgpu_autotune_export": (C.c_int, [C.c_char_p, C.c_int64, _i64p]),
            "sphgpu_autotune_import": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
            "sphgpu_graph_counts": (C.c_int, [vp, _i64p, _i64p]),
            "sphgpu_ensemble_begin": (C.c_int, [vp]),
            "sphgpu_ensemble_add": (C.c_int, [vp, C.POINTER(Member), C.POINTER(IntegratorParams), _i64p]),
            "sphgpu_ensemble_step": (C.c_int, [vp, C.POINTER(HydroParams), C.POINTER(IntegratorParams),
                                               C.POINTER(C.c_int)]),
            "sphgpu_ensemble_members": (C.c_int, [vp, C.c_int, _i64p, _dp, _dp, _i64p, _i64p, _i64p,
                                                  C.POINTER(C.c_int)]),
            "sphgpu_ensemble_failed": (C.c_int, [vp, C.c_int, _i64p, _i32p, _i64p, C.POINTER(C.c_int)]),
            "sphgpu_ensemble_set_dt_threshold": (C.c_int, [vp, C.c_double]),
            "sphgpu_ensemble_download": (C.c_int, [vp, C.c_int64, C.c_int, _dp]),
            "sphgpu_ensemble_upload": (C.c_int, [vp, C.c_int64, C.c_int, _dp]),
            "sphgpu_ensemble_remove": (C.c_int, [vp, C.c_int, _i64p]),
            "sphgpu_ensemble_set_clock": (C.c_int, [vp, C.c_int64, C.c_double, C.c_double, C.c_int64]),
            "sphgpu_ensemble_bytes_per_node": (C.c_int64, [C.c_int]),
            "sphgpu_serve": (C.c_int, [vp, C.POINTER(Request), C.c_int, C.POINTER(HydroParams),
                                       C.POINTER(IntegratorParams), C.POINTER(ServeParams)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct=_dp):
    return a.ctypes.data_as(ct)


def state_components(field, dim):
    if field in ("position", "velocity"):
        return dim
    if field == "H":
        return dim * (dim + 1) // 2
    return 1


def deriv_components(field, dim):
    i = DERIV_FIELDS.index(field) if field in DERIV_FIELDS else DERIV_ALIASES.index(field)
    a = DERIV_ALIASES[i]
    if a in ("DxDt", "DvDt", "XSPHDeltaV"):
        return dim
    if a in ("DvDx", "localDvDx", "M", "localM"):
        return dim * dim
    if a in ("DHDt", "Hideal", "massSecondMoment"):
        return dim * (dim + 1) // 2
    return 1


def autotune(on=True):
    """Turn the launch-shape autotuner on or off (process-wide)."""
    lib().sphgpu_autotune(1 if on else 0)


def autotune_table():
    """The autotuner's decisions as text (one workload class per line)."""
    L = lib()
    n = C.c_int64()
    L.sphgpu_autotune_export(None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    L.sphgpu_autotune_export(buf, n.value + 1, C.byref(n))
    return buf.value.decode()


def autotune_load(text):
    """Load autotuner decisions (the autotune_table format); returns the lines taken."""
    k = C.c_int()
    rc = lib().sphgpu_autotune_import(text.encode(), C.byref(k))
    if rc != 0:
        raise SphGpuError(rc, "bad autotune table")
    return k.value


class Sim:
    """One NodeList resident on one GPU (sphgpu_sim)."""

    def __init__(self, dim, device=0):
        self.L = lib()
        self.dim = dim
        h = C.c_void_p()
        rc = self.L.sphgpu_create(dim, device, C.byref(h))
        if rc != 0:
            raise SphGpuError(rc, "sphgpu_create failed (no HIP device?)")
        self.h = h
        self.n_internal = 0
        self.n_total = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.sphgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise SphGpuError(rc, self.L.sphgpu_last_error(self.h).decode())

    # kernel table ------------------------------------------------------------
    def set_table_bspline(self, npoints=1000):
        self._chk(self.L.sphgpu_set_table_bspline(self.h, npoints))
        self.npoints = npoints

    def set_table(self, npoints, extent, coeffs):
        c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1)
        assert c.size == 8 * npoints
        self._chk(self.L.sphgpu_set_table(self.h, npoints, extent, _p(c)))
        self.npoints = npoints

    def get_table(self):
        c = np.zeros(8 * self.npoints)
        self._chk(self.L.sphgpu_get_table(self.h, self.npoints, _p(c)))
        names = ["Ak", "Bk", "Ck", "Ag", "Bg", "Cg", "nperh", "wsum"]
        return {k: c[i * self.npoints:(i + 1) * self.npoints].copy() for i, k in enumerate(names)}

    def table_values(self, eta, Hdet=1.0):
        W = C.c_double()
        g = C.c_double()
        self._chk(self.L.sphgpu_table_values(self.h, eta, Hdet, C.byref(W), C.byref(g)))
        return W.value, g.value

    # state --------------------------------------------------------------------
    def set_nodes(self, n_internal, n_total):
        self._chk(self.L.sphgpu_set_nodes(self.h, n_internal, n_total))
        self.n_internal = n_internal
        self.n_total = n_total

    def upload(self, field, arr):
        """Upload n_total rows, or n_internal rows once the device made the ghosts."""
        fid = STATE_FIELDS[field]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        rows = a.shape[0]
        assert a.size == rows * state_components(field, self.dim), (field, a.shape)
        self._chk(self.L.sphgpu_upload(self.h, fid, _p(a), rows))

    def upload_state(self, state):
        names = {"pos": "position", "vel": "velocity", "mass": "mass", "rho": "mass density",
                 "eps": "specific thermal energy", "H": "H", "P": "pressure", "cs": "sound speed",
                 "omega": "grad h corrections", "fCl": "Cl multiplier", "fCq": "Cq multiplier",
                 "fshear": "Balsara shear correction"}
        for k, v in state.items():
            if k in names and v is not None:
                self.upload(names[k], v)

    def download(self, field, rows=None):
        c = state_components(field, self.dim)
        n = self.n_total if rows is None else rows
        a = np.zeros((n, c))
        self._chk(self.L.sphgpu_download(self.h, STATE_FIELDS[field], _p(a), n))
        return a[:, 0] if c == 1 else a

    # ghost boundaries ---------------------------------------------------------
    def set_boundaries(self, boundaries):
        arr = (Boundary * max(1, len(boundaries)))(*boundaries)
        self._chk(self.L.sphgpu_set_boundaries(self.h, len(boundaries), arr))
        self._sync_counts()

    def set_ghost_nodes(self, kernel_extent=2.0, cull=True):
        self._chk(self.L.sphgpu_set_ghost_nodes(self.h, kernel_extent, int(cull)))
        self._sync_counts()

    def apply_ghost_boundaries(self):
        self._chk(self.L.sphgpu_apply_ghost_boundaries(self.h))

    def _sync_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self._chk(self.L.sphgpu_node_counts(self.h, C.byref(a), C.byref(b)))
        self.n_internal, self.n_total = a.value, b.value

    def node_counts(self):
        self._sync_counts()
        return self.n_internal, self.n_total

    def ghost_controls(self):
        ng = self.n_total - self.n_internal
        c = np.zeros(max(ng, 1), dtype=np.int32)
        b = np.zeros(max(ng, 1), dtype=np.int32)
        self._chk(self.L.sphgpu_download_ghost_controls(self.h, _p(c, _i32p), _p(b, _i32p)))
        return c[:ng], b[:ng]

    # passes -------------------------------------------------------------------
    def set_row_order(self, order):
        self._chk(self.L.sphgpu_set_row_order(self.h, order))

    def build_neighbors(self, kernel_extent=2.0):
        self._chk(self.L.sphgpu_build_neighbors(self.h, kernel_extent))

    def neighbor_count(self):
        n = C.c_int64()
        self._chk(self.L.sphgpu_neighbor_count(self.h, C.byref(n)))
        return n.value

    def neighbors(self):
        total = self.neighbor_count()
        off = np.zeros(self.n_internal + 1, dtype=np.int64)
        nbr = np.zeros(max(total, 1), dtype=np.int32)
        self._chk(self.L.sphgpu_download_neighbors(self.h, _p(off, _i64p), _p(nbr, _i32p)))
        return off, nbr[:total]

    def set_precision(self, precision):
        """Arithmetic of sum_density (evaluate_derivatives takes params.precision)."""
        self._chk(self.L.sphgpu_set_precision(self.h, precision))

    def sum_density(self):
        self._chk(self.L.sphgpu_sum_density(self.h))

    def evaluate_derivatives(self, params=None, **kw):
        p = params if params is not None else hydro_params(**kw)
        self._chk(self.L.sphgpu_evaluate_derivatives(self.h, C.byref(p)))

    def derivative(self, field):
        i = DERIV_FIELDS.index(field) if field in DERIV_FIELDS else DERIV_ALIASES.index(field)
        c = deriv_components(field, self.dim)
        a = np.zeros((self.n_internal, c))
        self._chk(self.L.sphgpu_download_derivative(self.h, i, _p(a), self.n_internal))
        return a[:, 0] if c == 1 else a

    def derivatives(self):
        return {a: self.derivative(a) for a in DERIV_ALIASES}

    def pair_count(self):
        n = C.c_int64()
        self._chk(self.L.sphgpu_pair_count(self.h, C.byref(n)))
        return n.value

    def omega_gradh(self):
        """computeSPHOmegaGradhCorrection into the resident omega field."""
        self._chk(self.L.sphgpu_omega_gradh(self.h))

    # time integration ---------------------------------------------------------
    def eos_gamma_law(self, gamma, Pmin=-1.0e200, Pmax=1.0e200, floor=True):
        self._chk(self.L.sphgpu_eos_gamma_law(self.h, gamma, Pmin, Pmax, int(floor)))

    def hydro_dt(self, nPerh=1.51, cfl=0.25, use_velocity_magnitude=False):
        dt = C.c_double()
        self._chk(self.L.sphgpu_hydro_dt(self.h, nPerh, cfl, int(use_velocity_magnitude), C.byref(dt)))
        return dt.value

    def rk2_step(self, hydro=None, integ=None, max_time=FLT_MAX):
        hp = hydro if hydro is not None else hydro_params()
        ip = integ if integ is not None else integrator_params()
        dt = C.c_double()
        self._chk(self.L.sphgpu_rk2_step(self.h, C.byref(hp), C.byref(ip), max_time, C.byref(dt)))
        self._sync_counts()
        return dt.value

    def iterate_ideal_h(self, hydro=None, max_iterations=100, tolerance=1.0e-10, kernel_extent=2.0, cull=True):
        """iterateIdealH on the resident state; returns (iterations, last largest change)."""
        hp = hydro if hydro is not None else hydro_params()
        it, md = C.c_int(), C.c_double()
        self._chk(self.L.sphgpu_iterate_ideal_h(self.h, C.byref(hp), max_iterations, tolerance, kernel_extent,
                                                 int(cull), C.byref(it), C.byref(md)))
        self._sync_counts()
        return it.value, md.value

    def conservation(self, member=None):
        """Totals [mass, p(3), L(3), KE, TE] of the internal nodes (or one ensemble member)."""
        out = np.zeros(9)
        if member is None:
            self._chk(self.L.sphgpu_conservation(self.h, _p(out)))
        else:
            self._chk(self.L.sphgpu_ensemble_conservation(self.h, member, _p(out)))
        return out

    def dt_inputs(self, first=0, rows=None):
        """The dt vote's inputs of rows [first, first + rows): (rows, 2 + dim)."""
        n = (self.node_counts()[0] - first) if rows is None else rows
        a = np.zeros((n, 2 + self.dim))
        self._chk(self.L.sphgpu_get_dt_inputs(self.h, first, n, _p(a)))
        return a

    def set_dt_inputs(self, a, first=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.L.sphgpu_set_dt_inputs(self.h, first, a.shape[0], _p(a)))

    def clock(self):
        t, ldt, cyc = C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.L.sphgpu_get_time(self.h, C.byref(t), C.byref(ldt), C.byref(cyc)))
        return t.value, ldt.value, cyc.value

    def set_clock(self, t=0.0, last_dt=1.0e-5, cycle=0):
        self._chk(self.L.sphgpu_set_time(self.h, t, last_dt, cycle))

    def timings(self):
        t = np.zeros(6)
        self._chk(self.L.sphgpu_last_timings(self.h, _p(t)))
        return dict(neighbors_ms=t[0], density_ms=t[1], derivs_ms=t[2], pack_ms=t[3], ghosts_ms=t[4],
                    cull_ms=t[5])

    def synchronize(self):
        self._chk(self.L.sphgpu_synchronize(self.h))

    def set_node_lists(self, counts, gamma=None, sum_all=True):
        """Several NodeLists: NodeList L owns the next counts[L] internal rows;
        gamma (one per NodeList, optional) sets each its gamma-law EOS;
        sum_all = sumMassDensityOverAllNodeLists."""
        c = np.ascontiguousarray(counts, dtype=np.int64)
        g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64)
        self._chk(self.L.sphgpu_set_node_lists(self.h, len(c), _p(c, _i64p), _p(g) if g is not None else None,
                                               1 if sum_all else 0))

    def set_graphs(self, on=True):
        """Replay the device part of each time step from a captured HIP graph."""
        self._chk(self.L.sphgpu_set_graphs(self.h, 1 if on else 0))

    def graph_counts(self):
        """(steps captured as graphs, steps replayed from a graph)."""
        c, r = C.c_int64(), C.c_int64()
        self._chk(self.L.sphgpu_graph_counts(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def conservation_noalloc(self, first, rows, out_device_ptr):
        """sphgpu_conservation_noalloc: the totals of rows [first, first + rows)
        into 9 doubles of device memory at out_device_ptr (an integer address),
        asynchronously on the sim's stream."""
        self._chk(self.L.sphgpu_conservation_noalloc(self.h, int(first), int(rows), C.c_void_p(int(out_device_ptr))))

    # ensembles -------------------------------------------------------------------
    def ensemble_begin(self):
        """Switch this (empty) sim to an ensemble of independent members."""
        self._chk(self.L.sphgpu_ensemble_begin(self.h))
        self._members = {}

    def ensemble_add(self, state, gamma, t_end, integ=None):
        """Append a member from a state dict (pos, vel, mass, rho, eps, H);
        returns its id."""
        ip = integ if integ is not None else integrator_params(gamma=gamma)
        f = lambda k: np.ascontiguousarray(state[k], dtype=np.float64)
        arrs = {k: f(k) for k in ("pos", "vel", "mass", "rho", "eps", "H")}
        n = arrs["mass"].shape[0]
        m = Member(n=n, gamma=gamma, tEnd=t_end, **{k: _p(v) for k, v in arrs.items()})
        mid = C.c_int64()
        self._chk(self.L.sphgpu_ensemble_add(self.h, C.byref(m), C.byref(ip), C.byref(mid)))
        self._members[mid.value] = n
        return mid.value

    def ensemble_step(self, hydro=None, integ=None):
        """One RK2 step of every member; returns how many reached their end time."""
        hp = hydro if hydro is not None else hydro_params()
        ip = integ if integ is not None else integrator_params()
        done = C.c_int()
        self._chk(self.L.sphgpu_ensemble_step(self.h, C.byref(hp), C.byref(ip), C.byref(done)))
        return done.value

    def ensemble_failed(self):
        """Members the last ensemble step evicted: list of dicts (id, reason, cycle)."""
        cnt = C.c_int()
        self._chk(self.L.sphgpu_ensemble_failed(self.h, 0, None, None, None, C.byref(cnt)))
        n = cnt.value
        ids, cyc = np.zeros(max(n, 1), dtype=np.int64), np.zeros(max(n, 1), dtype=np.int64)
        why = np.zeros(max(n, 1), dtype=np.int32)
        self._chk(self.L.sphgpu_ensemble_failed(self.h, n, _p(ids, _i64p), _p(why, _i32p), _p(cyc, _i64p),
                                                C.byref(cnt)))
        for q in range(n):
            self._members.pop(int(ids[q]), None)
        return [dict(id=int(ids[q]), reason=int(why[q]), cycle=int(cyc[q])) for q in range(n)]

    def ensemble_set_dt_threshold(self, dt_threshold):
        self._chk(self.L.sphgpu_ensemble_set_dt_threshold(self.h, float(dt_threshold)))

    def ensemble_members(self):
        """List of dicts (id, t, last_dt, cycle, first, rows) in row order."""
        cnt = C.c_int()
        self._chk(self.L.sphgpu_ensemble_members(self.h, 0, None, None, None, None, None, None, C.byref(cnt)))
        n = cnt.value
        ids, cyc, first, rows = (np.zeros(max(n, 1), dtype=np.int64) for _ in range(4))
        t, ldt = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        self._chk(self.L.sphgpu_ensemble_members(self.h, n, _p(ids, _i64p), _p(t), _p(ldt), _p(cyc, _i64p),
                                                 _p(first, _i64p), _p(rows, _i64p), C.byref(cnt)))
        return [dict(id=int(ids[q]), t=float(t[q]), last_dt=float(ldt[q]), cycle=int(cyc[q]), first=int(first[q]),
                     rows=int(rows[q])) for q in range(n)]

    def ensemble_download(self, member_id, field):
        c = state_components(field, self.dim)
        a = np.zeros((self._members[member_id], c))
        self._chk(self.L.sphgpu_ensemble_download(self.h, member_id, STATE_FIELDS[field], _p(a)))
        return a[:, 0] if c == 1 else a

    def ensemble_set_clock(self, member_id, t, last_dt, cycle):
        self._chk(self.L.sphgpu_ensemble_set_clock(self.h, member_id, t, last_dt, cycle))

    def ensemble_upload(self, member_id, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self._members[member_id] * state_components(field, self.dim), (field, a.shape)
        self._chk(self.L.sphgpu_ensemble_upload(self.h, member_id, STATE_FIELDS[field], _p(a)))

    def ensemble_remove(self, ids):
        ids = np.ascontiguousarray(list(ids), dtype=np.int64)
        self._chk(self.L.sphgpu_ensemble_remove(self.h, len(ids), _p(ids, _i64p)))
        for q in ids:
            self._members.pop(int(q), None)

    def serve(self, requests, hydro=None, integ=None, budget=0, max_members=1024, ghost_factor=1.5, dt_threshold=0.0):
        """Run the request loop (sphgpu_serve) on this empty ensemble.
        requests: dicts with state, gamma, t_end, arrival (seconds).  Returns
        per request: admitted, finished, steps, status (MEMBER_*) and the final
        pos, vel, rho, eps, H."""
        hp = hydro if hydro is not None else hydro_params()
        ip = integ if integ is not None else integrator_params()
        dim, ns = self.dim, self.dim * (self.dim + 1) // 2
        keep, outs = [], []
        reqs = (Request * max(1, len(requests)))()
        for q, r in enumerate(requests):
            st = r["state"]
            arrs = {k: np.ascontiguousarray(st[k], dtype=np.float64) for k in ("pos", "vel", "mass", "rho", "eps", "H")}
            n = arrs["mass"].shape[0]
            o = dict(pos=np.zeros((n, dim)), vel=np.zeros((n, dim)), rho=np.zeros(n), eps=np.zeros(n),
                     H=np.zeros((n, ns)))
            keep.append(arrs)
            outs.append(o)
            reqs[q].member = Member(n=n, gamma=r["gamma"], tEnd=r["t_end"], **{k: _p(v) for k, v in arrs.items()})
            reqs[q].arrival = r.get("arrival", 0.0)
            reqs[q].outPos, reqs[q].outVel = _p(o["pos"]), _p(o["vel"])
            reqs[q].outRho, reqs[q].outEps, reqs[q].outH = _p(o["rho"]), _p(o["eps"]), _p(o["H"])
        sp = ServeParams(hbmBudget=int(budget), maxMembers=int(max_members), ghostFactor=float(ghost_factor),
                         dtThreshold=float(dt_threshold))
        self._chk(self.L.sphgpu_serve(self.h, reqs, len(requests), C.byref(hp), C.byref(ip), C.byref(sp)))
        res = []
        for q, o in enumerate(outs):
            o.update(admitted=reqs[q].admitted, finished=reqs[q].finished, steps=int(reqs[q].steps),
                     status=int(reqs[q].status))
            res.append(o)
        return res

    def copy_probe(self, nbytes, repeats=5):
        """Average time (ms) of a plain device-to-device streaming copy kernel."""
        ms = C.c_double()
        self._chk(self.L.sphgpu_copy_probe(self.h, nbytes, repeats, C.byref(ms)))
        return ms.value

    def gather_probe(self, loads, repeats=5):
        """Average time (ms) of a kernel that walks the current neighbour rows
        as the pair and density kernels do and only loads `loads` (2, 3 or 6)
        16-B pieces of each neighbour's record: their gather ceiling."""
        ms = C.c_double()
        self._chk(self.L.sphgpu_gather_probe(self.h, loads, repeats, C.byref(ms)))
        return ms.value
