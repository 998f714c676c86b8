"""Extract the Spada et al. (2011) disc-load benchmark curves shipped with the
reference (data/benchmarks/disc_spada2011/{u,n}_disc.txt, see
data/benchmarks/PROVENANCE.md there) into a small JSON fixture. The files hold
201 rows (colatitude 0:0.1:20 deg, ascending) by 6 columns (t = 0, 1, 2, 5,
10, 100 kyr; column 1 is the elastic response) of vertical displacement u
and geoid n [m] for a 10-degree disc of 1000 m ice (rho 931) on M3-L70-V01.
Only every tenth row is data (the curves are piecewise linear on a 1-degree
grid, PROVENANCE.md "Sampling"), so the fixture keeps the 21 rows at whole
degrees.

Run once in the build container: python tests/golden/make_disc_fixture.py
"""
import json
import os

SRC = "/root/reference/data/benchmarks/disc_spada2011/"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spada2011_disc.json")


def main():
    out = {"source": "reference data/benchmarks/disc_spada2011/{u,n}_disc.txt (rows at whole degrees)",
           "theta_deg": list(range(21)), "t_kyr": [0, 1, 2, 5, 10, 100]}
    for name in ("u", "n"):
        rows = [[float(v) for v in l.split()] for l in open(SRC + name + "_disc.txt") if l.strip()]
        assert len(rows) == 201 and all(len(r) == 6 for r in rows)
        out[name] = [rows[10 * d] for d in range(21)]
    json.dump(out, open(OUT, "w"), indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
