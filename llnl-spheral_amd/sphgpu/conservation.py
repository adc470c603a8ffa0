"""Conservation history in the reference's `.conserve` column format.

SpheralConservation (src/SimulationControl/SpheralConservation.py:22-200)
keeps, per recorded cycle, the total mass, linear and angular momentum, the
kinetic, thermal and package energies, and writes them as a '#' header of
'%14s ' quoted labels followed by one row of '%14.8g ' values per cycle.  The
totals here come from the device (sphgpu_conservation); the SPH hydro package
carries no package energy (Physics::extraEnergy is 0), so "Pkg E" is 0.
"""
import math

LABELS = ['"cycle"', '"time"', '"Mass"',
          '"Lin Mom Mag"', '"Lin Mom X"', '"Lin Mom Y"', '"Lin Mom Z"',
          '"Ang Mom Mag"', '"Ang Mom X"', '"Ang Mom Y"', '"Ang Mom Z"',
          '"Total E"', '"Kin E"', '"Therm E"', '"Pkg E"']


class ConservationHistory:
    """History of one simulation (or one ensemble member)."""

    def __init__(self, sim, member=None):
        self.sim = sim
        self.member = member
        self.rows = []
        self.update(0, 0.0)

    def update(self, cycle, time):
        t = self.sim.conservation(self.member)
        mass, p, L, ke, te = t[0], t[1:4], t[4:7], t[7], t[8]
        pkg = 0.0
        mag = lambda v: math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        self.rows.append([cycle, time, mass, mag(p), p[0], p[1], p[2], mag(L), L[0], L[1], L[2],
                          ke + te + pkg, ke, te, pkg])

    def write(self, filename):
        with open(filename, "w") as f:
            f.write("#")
            for lab in LABELS:
                f.write("%14s " % lab)
            f.write("\n")
            for r in self.rows:
                for v in r:
                    f.write("%14.8g " % v)
                f.write("\n")


def read(filename):
    """Rows of a .conserve file as lists of floats (header skipped)."""
    out = []
    with open(filename) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            out.append([float(x) for x in line.split()])
    return out
