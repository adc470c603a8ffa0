"""Restart snapshots of a simulation or of one ensemble member.

The reference's restart files (RestartRegistrar, src/DataOutput/RestartRegistrar.hh:28-66)
hold every registered object's dumpState: the NodeList fields and the
integrator's lastDt, currentTime and currentCycle
(src/Integrator/Integrator.cc:794-823).  A snapshot here holds the same state
of the internal nodes (position, velocity, mass, mass density, specific
thermal energy, H, the grad-h correction, pressure and sound speed, which the
last update set before the finalize's density sum), the clock and the
derivatives the next dt vote reads (the reference dumps pressure, sound speed
and its derivative fields too, src/SPH/SPHHydroBase.cc:1283-1318), in a numpy
.npz file (loaded without pickles); ghosts and neighbours are remade by the
next step.
"""
import numpy as np

FIELDS = {"pos": "position", "vel": "velocity", "mass": "mass", "rho": "mass density",
          "eps": "specific thermal energy", "H": "H", "omega": "grad h corrections",
          "P": "pressure", "cs": "sound speed"}


def save(sim, path, member=None, gamma=None, t_end=None):
    """Write the internal-node state and clock of `sim` (or of ensemble member
    `member`, with its gamma and end time) to `path`."""
    out = {"dim": np.int64(sim.dim)}
    if member is None:
        n = sim.node_counts()[0]
        for k, f in FIELDS.items():
            out[k] = sim.download(f, rows=n)
        out["dtin"] = sim.dt_inputs(0, n)
        t, ldt, cyc = sim.clock()
    else:
        for k, f in FIELDS.items():
            out[k] = sim.ensemble_download(member, f)
        mb = [m for m in sim.ensemble_members() if m["id"] == member][0]
        out["dtin"] = sim.dt_inputs(mb["first"], mb["rows"])
        t, ldt, cyc = mb["t"], mb["last_dt"], mb["cycle"]
        out["gamma"] = np.float64(gamma)
        out["t_end"] = np.float64(t_end)
    out["clock"] = np.array([t, ldt, float(cyc)])
    np.savez(path, **out)


def load(path):
    """The snapshot as a dict (state arrays, clock, optional gamma/t_end)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def restore(sim, path):
    """Restore a plain simulation: n internal nodes, state and clock."""
    z = load(path)
    n = z["mass"].shape[0]
    sim.set_nodes(n, n)
    sim.upload_state({k: z[k] for k in FIELDS})
    t, ldt, cyc = z["clock"]
    sim.set_clock(float(t), float(ldt), int(cyc))
    if int(cyc) > 0:
        sim.set_dt_inputs(z["dtin"], 0)


def restore_member(sim, path, integ=None):
    """Add a snapshot to an ensemble as a member with its clock; returns its id."""
    z = load(path)
    mid = sim.ensemble_add({k: z[k] for k in FIELDS}, float(z["gamma"]), float(z["t_end"]), integ)
    for k in ("P", "cs", "omega"):
        sim.ensemble_upload(mid, FIELDS[k], z[k])
    t, ldt, cyc = z["clock"]
    sim.ensemble_set_clock(mid, float(t), float(ldt), int(cyc))
    if int(cyc) > 0:
        mb = [m for m in sim.ensemble_members() if m["id"] == mid][0]
        sim.set_dt_inputs(z["dtin"], mb["first"])
    return mid
