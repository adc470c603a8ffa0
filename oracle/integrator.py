"""Test infrastructure: CPU oracle for the SPH time step (FP64).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It restates, in the reference's order of operations:

  * SynchronousRK2::step                       src/Integrator/SynchronousRK2.cc:78-144
      preStepInitialize (setGhostNodes + applyGhostBoundaries)
                                               src/Integrator/Integrator.cc:221-252, 443-650
      selectDt (dtGrowth, dtMin/dtMax clamps)  src/Integrator/Integrator.cc:167-212
      state update to t + dt/2, ghosts, second evaluateDerivatives, update from
      the saved state with the midpoint derivatives, ghosts, finalize
      (SPHHydroBase::finalize sum density)     src/SPH/SPHHydroBase.cc:965-992
      enforceBoundaries                        src/Integrator/Integrator.cc:726-738
  * State::update policies registered by SPHHydroBase::registerState
                                               src/SPH/SPHHydroBase.cc:238-336
      position, velocity, specific energy: f += multiplier*df
                                               src/DataBase/IncrementFieldList.cc:107-151
      mass density: min(rhoMax, max(rhoMin, rho + multiplier*DrhoDt))
                                               src/DataBase/IncrementBoundedState.cc:120-136
      H: ideal H bounded by 1/hmax, 1/hmin (ReplaceBoundedState, :121-137); the
         ideal H of the SPH pass is h1^-1 I, so the eigenvalue bound is a clamp
      pressure, sound speed: GammaLawGas       src/Material/GammaLawGas.cc:163-223
  * GenericHydro::dt                           src/Physics/GenericHydro.cc:114-331

The arithmetic lives in the C oracle (ghosts, connectivity, sum density,
evaluateDerivatives, eigenvalues, boundary enforcement); this module only
sequences it and applies the elementwise updates, which numpy evaluates in
IEEE double precision without contraction.
"""
import numpy as np

import oracle

FLT_MIN = 1.1754943508222875e-38
FLT_MAX = 3.4028234663852886e38


def gamma_law_eos(rho, eps, gamma, Pmin=-1.0e200, Pmax=1.0e200):
    """GammaLawGas::pressure / soundSpeed with applyPressureLimits (floor type)."""
    P, cs = oracle.gamma_law(rho, eps, gamma, Pmin, Pmax, floor=True)
    return P, cs


def hydro_dt(dim, n, full, prev, off, nbr, nPerh, cfl, use_velocity_magnitude=False):
    """GenericHydro::dt over the internal nodes: cfl times the smallest of the
    sound-speed, viscous, divergence, pairwise-velocity and acceleration limits."""
    H = full["H"][:n]
    lam = oracle.max_eigen_n(dim, H)
    nodeScale = 1.0 / lam / nPerh
    cs = full["cs"][:n]
    rho = full["rho"][:n]
    cand = [nodeScale / (cs + FLT_MIN)]
    csq = np.sqrt(prev["maxViscousPressure"] / rho)
    cand.append(nodeScale / (csq + FLT_MIN))
    DvDx = prev["DvDx"].reshape(n, dim, dim)
    divv = DvDx[:, 0, 0]
    for a in range(1, dim):
        divv = divv + DvDx[:, a, a]
    cand.append(1.0 / (np.abs(divv) + FLT_MIN))
    vel = full["vel"].reshape(-1, dim)
    rows = np.repeat(np.arange(n), np.diff(off))
    d = vel[rows] - vel[nbr]
    if dim == 1:
        vij = np.abs(d[:, 0])
    else:
        s = d[:, 0] * d[:, 0]
        for a in range(1, dim):
            s = s + d[:, a] * d[:, a]
        vij = np.sqrt(s)
    pair = nodeScale[rows] * (1.0 / np.maximum(1.0e-30, vij))
    pmin = np.full(n, FLT_MAX)
    np.minimum.at(pmin, rows, pair)
    cand.append(pmin)
    acc = prev["DvDt"].reshape(n, dim)
    if dim == 1:
        amag = np.abs(acc[:, 0])
    else:
        s = acc[:, 0] * acc[:, 0]
        for a in range(1, dim):
            s = s + acc[:, a] * acc[:, a]
        amag = np.sqrt(s)
    cand.append(np.sqrt(nodeScale / (amag + FLT_MIN)))
    if use_velocity_magnitude:
        v = vel[:n]
        s = v[:, 0] * v[:, 0]
        for a in range(1, dim):
            s = s + v[:, a] * v[:, a]
        cand.append(nodeScale / ((np.abs(v[:, 0]) if dim == 1 else np.sqrt(s)) + 1.0e-10))
    minDt = FLT_MAX
    for c in cand:
        if c.size:
            minDt = min(minDt, float(c.min()))
    return cfl * minDt


class RK2:
    """One NodeList advanced by SynchronousRK2 with the SPH pass.

    state: internal rows (pos, vel, mass, rho, eps, H, optional omega);
    P and cs are set from the gamma-law EOS at construction."""

    def __init__(self, dim, state, boundaries, gamma, params=None, table_points=1000, cfl=0.25, dtMin=0.0,
                 dtMax=FLT_MAX, dtGrowth=2.0, rhoMin=1.0e-10, rhoMax=1.0e10, hmin=1.0e-20, hmax=1.0e20,
                 ext=2.0, cull=True, Pmin=-1.0e200, Pmax=1.0e200, gradh=False, compatible=False,
                 lists=None, gammas=None, sum_all=True):
        """lists: NodeList of every internal row (several NodeLists, each a
        contiguous block in NodeList order), gammas: each NodeList's gamma,
        sum_all: sumMassDensityOverAllNodeLists."""
        self.dim = dim
        self.lists = None if lists is None else np.asarray(lists, dtype=np.int32)
        self.gammas = gammas
        self.sum_all = sum_all
        self.n = state["pos"].shape[0]
        self.st = {k: np.array(v, dtype=np.float64) for k, v in state.items() if v is not None}
        self.st.setdefault("omega", np.ones(self.n))
        self.b = list(boundaries)
        self.gamma = gamma
        self.p = params if params is not None else oracle.make_params()
        self.T = oracle.Table(dim, table_points)
        self.cfl, self.dtMin, self.dtMax, self.dtGrowth = cfl, dtMin, dtMax, dtGrowth
        self.rhoMin, self.rhoMax = rhoMin, rhoMax
        self.hmaxInv, self.hminInv = 1.0 / hmax, 1.0 / hmin
        self.ext, self.cull = ext, cull
        self.Pmin, self.Pmax = Pmin, Pmax
        self.gradh = gradh
        self.compatible = compatible
        self.st["P"], self.st["cs"] = self._eos(self.st["rho"], self.st["eps"])
        self.t = 0.0
        self.lastDt = 1.0e-5
        self.cycle = 0
        nsym = dim * (dim + 1) // 2
        self.prev = {"maxViscousPressure": np.zeros(self.n), "DvDx": np.zeros((self.n, dim * dim)),
                     "DvDt": np.zeros((self.n, dim))}
        self.nsym = nsym

    # -- pieces ----------------------------------------------------------------
    def _eos(self, rho, eps):
        """GammaLawGas of the internal rows (each NodeList its own gamma)."""
        if self.lists is None or self.gammas is None:
            return gamma_law_eos(rho, eps, self.gamma, self.Pmin, self.Pmax)
        P, cs = np.zeros_like(rho), np.zeros_like(rho)
        for L, g in enumerate(self.gammas):
            sel = self.lists == L
            P[sel], cs[sel] = gamma_law_eos(rho[sel], eps[sel], g, self.Pmin, self.Pmax)
        return P, cs

    def _ghosts(self):
        if self.b:
            full, off, nbr, info = oracle.ghost_pipeline(self.dim, self.st, self.b, self.ext, self.cull,
                                                         lists=self.lists)
        else:
            full = {k: v.copy() for k, v in self.st.items()}
            off, nbr = oracle.connectivity(self.dim, self.n, full["pos"], full["H"], self.ext)
            info = None
            if self.lists is not None:
                full["list"] = self.lists.copy()
        return full, off, nbr, info

    def _apply_ghosts(self, full, info):
        """Integrator::applyGhostBoundaries: positions and H of the ghosts, then
        every package field, boundary by boundary."""
        if info is None:
            return
        n = self.n
        c, so = info["control"], info["sub_of"]
        full["pos"] = oracle.apply_ghosts(self.dim, n, self.b, c, so, full["pos"], 3)
        full["H"] = oracle.apply_ghosts(self.dim, n, self.b, c, so, full["H"], 2)
        full["vel"] = oracle.apply_ghosts(self.dim, n, self.b, c, so, full["vel"], 1)
        for k in ("mass", "rho", "eps", "P", "cs", "omega"):
            full[k] = oracle.apply_ghosts(self.dim, n, self.b, c, so, full[k], 0)

    def _update(self, full, base, d, mult):
        """State::update on the internal rows of `full` from `base` (the
        current state or the saved state0)."""
        n, dim = self.n, self.dim
        if self.compatible:
            eps = self._compatible_eps(full, base, d, mult)
        else:
            eps = base["eps"] + mult * d["DepsDt"]
        full["pos"][:n] = base["pos"] + mult * d["DxDt"].reshape(n, dim)
        full["vel"][:n] = base["vel"] + mult * d["DvDt"].reshape(n, dim)
        full["eps"][:n] = eps
        full["rho"][:n] = np.minimum(self.rhoMax, np.maximum(self.rhoMin, base["rho"] + mult * d["DrhoDt"]))
        hid = d["Hideal"].reshape(n, self.nsym)
        full["H"][:n] = np.minimum(self.hminInv, np.maximum(self.hmaxInv, hid)) * (hid != 0.0) + hid * (hid == 0.0)
        full["P"][:n], full["cs"][:n] = self._eos(full["rho"][:n], full["eps"][:n])

    def _derivs(self, full, off, nbr, info=None):
        if self.gradh:
            # SPHHydroBase::initialize: grad-h correction, then its ghost boundaries
            full["omega"][:self.n] = oracle.omega_gradh(self.T, self.n, full["pos"], full["H"], off, nbr,
                                                        lists=full.get("list"))
            if info is not None:
                full["omega"] = oracle.apply_ghosts(self.dim, self.n, self.b, info["control"], info["sub_of"],
                                                    full["omega"], 0)
        d = oracle.eval_derivs(self.T, self.p, self.n, full, off, nbr, pair_acc=self.compatible,
                               lists=full.get("list"))
        if self.compatible:
            # SPHHydroBase::finalizeDerivatives: boundaries on DvDt and DepsDt
            n, dim = self.n, self.dim
            nt = full["pos"].shape[0]
            acc = np.zeros((nt, dim))
            acc[:n] = d["DvDt"].reshape(n, dim)
            de0 = np.zeros(nt)
            de0[:n] = d["DepsDt"]
            if info is not None:
                acc = oracle.apply_ghosts(dim, n, self.b, info["control"], info["sub_of"], acc, 1)
                de0 = oracle.apply_ghosts(dim, n, self.b, info["control"], info["sub_of"], de0, 0)
            d["_acc_all"], d["_depsdt0_all"] = acc, de0
            # the rows the pair accelerations sit in (several NodeLists: the
            # delivery order, whose owned entries come in ConnectivityMap
            # order, as the policy walks them)
            d["_rows"] = (off, d.get("nbr_order", nbr))
        return d

    def _compatible_eps(self, full, base, d, mult):
        """SpecificThermalEnergyPolicy: eps from the pair accelerations, with
        the velocities the state holds before the update."""
        off, nbr = d["_rows"]
        de = oracle.compatible_depsdt(self.dim, self.n, off, nbr, full["mass"], full["vel"], d["_acc_all"],
                                      d["_depsdt0_all"], d["pairAcc"], mult)
        return base["eps"] + de * mult

    # -- the step ----------------------------------------------------------------
    def step(self, maxTime=FLT_MAX):
        n, dim = self.n, self.dim
        t = self.t
        full, off, nbr, info = self._ghosts()
        dtMin = min(self.dtMin, maxTime - t)
        dtMax = min(self.dtMax, maxTime - t)
        vote = hydro_dt(dim, n, full, self.prev, off, nbr, self.p.nPerh, self.cfl)
        dt = dtMax
        if vote > 0.0 and vote < dt:
            dt = vote
        dt = min(dt, self.dtGrowth * self.lastDt)
        dt = min(dtMax, max(dtMin, dt))
        hdt = 0.5 * dt
        d1 = self._derivs(full, off, nbr, info)
        # State state0(state): every registered field, ghost rows included
        state0 = {k: full[k].copy() for k in ("pos", "vel", "mass", "eps", "rho", "H", "P", "cs", "omega")}
        self._update(full, {k: v[:n].copy() for k, v in state0.items()}, d1, hdt)
        self._apply_ghosts(full, info)
        d2 = self._derivs(full, off, nbr, info)
        for k, v in state0.items():       # state.assign(state0)
            full[k][:] = v
        self._update(full, {k: v[:n] for k, v in state0.items()}, d2, dt)
        self._apply_ghosts(full, info)
        # SPHHydroBase::finalize: sum density with the step's connectivity
        full["rho"][:n] = oracle.sum_density(self.T, n, full["pos"], full["mass"], full["H"], off, nbr,
                                             lists=full.get("list"), sum_all=self.sum_all)
        for k in ("pos", "vel", "mass", "rho", "eps", "H", "P", "cs", "omega"):
            self.st[k] = full[k][:n].copy()
        if self.b:
            oracle.enforce_boundaries(dim, n, self.b, self.st["pos"], self.st["vel"], self.st["H"])
        self.prev = {"maxViscousPressure": d2["maxViscousPressure"].copy(), "DvDx": d2["DvDx"].reshape(n, -1).copy(),
                     "DvDt": d2["DvDt"].reshape(n, -1).copy()}
        self.cycle += 1
        self.t = t + dt
        self.lastDt = dt
        return dt


def iterate_ideal_h(dim, state, boundaries, max_iterations=100, tolerance=1.0e-10, params=None, table_points=1000,
                    ext=2.0, cull=True, lists=None):
    """iterateIdealH (src/Utilities/iterateIdealH.cc:30-310) with the SPH
    smoothing scale, nPerhForIteration = 0, no spherical start, no fixed
    determinant: each iteration rebuilds the ghosts and the connectivity from
    the current H, then passes over the internal nodes (oracle.ideal_h_pass);
    the loop ends after max_iterations or when the largest change is within
    the tolerance.  lists (NodeList of every internal node): several
    NodeLists, each making its own ghosts; neighbours of another NodeList are
    weighted by m_j rho_i / (m_i rho_j).  Returns (H of the internal nodes,
    iterations, last largest change)."""
    n = state["pos"].shape[0]
    T = oracle.Table(dim, table_points)
    p = params if params is not None else oracle.make_params()
    st = {k: np.array(state[k], dtype=np.float64) for k in ("pos", "H", "mass", "rho")}
    deltaH = np.full(n, 2.0 * tolerance)
    maxDelta = 2.0 * tolerance
    it = 0
    while it < max_iterations and maxDelta > tolerance:
        it += 1
        if boundaries:
            full, off, nbr, _ = oracle.ghost_pipeline(dim, st, boundaries, ext, cull, lists=lists)
        else:
            full = dict(st)
            off, nbr = oracle.connectivity(dim, n, st["pos"], st["H"], ext)
            if lists is not None:
                full["list"] = np.asarray(lists, dtype=np.int32)
        H = np.ascontiguousarray(full["H"], dtype=np.float64).copy()
        if lists is None:
            maxDelta = oracle.ideal_h_pass(T, p, n, full["pos"], H, deltaH, off, nbr, tolerance)
        else:
            maxDelta = oracle.ideal_h_pass(T, p, n, full["pos"], H, deltaH, off, nbr, tolerance, lists=full["list"],
                                           mass=full["mass"], rho=full["rho"])
        st["H"] = H[:n].copy()
    return st["H"], it, maxDelta
