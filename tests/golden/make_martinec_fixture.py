"""Extract the Martinec et al. (2018) sea-level benchmark profiles shipped
with the reference (data/benchmarks/sle_martinec2018/*_SBK.dat, giapy
output of the benchmark; reader: tests/test_benchmark_sle.f90 read_fig,
tests/test_benchmark_martinec.f90 read_ref) into one compressed fixture.

Per SLE case C2 / D3 / E2 / F1 and figure fig10..fig13 the array
`<case>_<fig>` holds the rows (col1, u, -N*g, sea surface, sea level):
col1 is colatitude [deg] (fig10, fig12) or longitude [deg] (fig11, fig13),
u the vertical displacement [m], -N*g the potential increment [m^2 s^-2],
the sea surface N + esl and the sea level rsl [m]. The horizontal columns
are not kept. `A_fig10` holds (colatitude, uplift, geoid) of the pure
loading case A.

Run once in the build container: python tests/golden/make_martinec_fixture.py
"""
import os

import numpy as np

SRC = "/root/reference/data/benchmarks/sle_martinec2018"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "martinec2018.npz")


def _rows(path, ncol):
    out = []
    for line in open(path):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        v = [float(x) for x in s.split()]
        if len(v) == ncol:
            out.append(v)
    return np.array(out)


def main():
    arrs = {}
    for case in ("C2", "D3", "E2", "F1"):
        for fig in ("fig10", "fig11", "fig12", "fig13"):
            a = _rows(os.path.join(SRC, "%s_%s_SBK.dat" % (case, fig)), 7)
            arrs["%s_%s" % (case, fig)] = a[:, [0, 1, 4, 5, 6]]
    a = _rows(os.path.join(SRC, "A_fig10_SBK.dat"), 4)
    arrs["A_fig10"] = a[:, [0, 1, 3]]
    np.savez_compressed(OUT, **arrs)
    print("wrote", OUT, {k: v.shape for k, v in arrs.items()})


if __name__ == "__main__":
    main()
