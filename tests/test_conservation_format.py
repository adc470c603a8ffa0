"""The .conserve history format of SpheralConservation
(src/SimulationControl/SpheralConservation.py:166-200): a '#' header of
'%14s ' quoted labels, then '%14.8g ' per value.  The value text is pinned
against the first rows of the reference's own history file
tests/Hydro/Noh/Results/Noh-planar-1d.conserve (kept as the fixture
tests/golden/noh_planar_1d_conserve_head.txt).  That file was written by an
older writer (no "Pkg E" column, one space less between columns), so the
first 14 values are compared as text tokens and the column width is checked
against the current writer's format."""
import os

from sphgpu import conservation

HERE = os.path.dirname(os.path.abspath(__file__))


class FakeSim:
    def __init__(self, rows):
        self.rows = list(rows)

    def conservation(self, member=None):
        return self.rows.pop(0)


def test_rows_match_reference_file(tmp_path):
    ref = [l for l in open(os.path.join(HERE, "golden", "noh_planar_1d_conserve_head.txt")) if not l.startswith("#")]
    vals = conservation.read(os.path.join(HERE, "golden", "noh_planar_1d_conserve_head.txt"))
    # totals that reproduce the reference rows: [mass, p(3), L(3), KE, TE]
    totals = [[r[2], r[4], r[5], r[6], r[8], r[9], r[10], r[12], r[13]] for r in vals]
    h = conservation.ConservationHistory(FakeSim(totals))
    for r in vals[1:]:
        h.update(int(r[0]), r[1])
    out = tmp_path / "x.conserve"
    h.write(str(out))
    lines = open(out).read().splitlines()
    assert lines[0].startswith('#       "cycle" ')
    assert len(lines[0].split('"')) == 2 * len(conservation.LABELS) + 1
    for got, want in zip(lines[1:], ref):
        g, w = got.split(), want.split()
        # "Total E" is the sum of the unrounded energies in the reference run;
        # here it is the sum of their 8-digit prints
        assert g[:11] + g[12:14] == w[:11] + w[12:14], (got, want)
        assert abs(float(g[11]) - float(w[11])) <= 2e-8 * abs(float(w[11]))
        assert got.split()[14] == "0"          # Pkg E
        assert len(got) == 15 * 15 and all(got[15 * q + 14] == " " for q in range(15))
