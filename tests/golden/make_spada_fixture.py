"""Extract the Spada et al. (2011) GIA benchmark Test 3/2 table (polar motion
under a Heaviside load, Table 14) shipped with the reference
(data/benchmarks/rotation_spada2011/polar_motion_test3-2.txt) into a small JSON
fixture: per load (cap, disc, point), case (incl / excl Chandler wobble) and
time [kyr], |m| [deg] of the two contributors (Vb, Gs), plus the published
geometrical factors G_sigma (Table 14 caption) from the file's header.

Run once in the build container: python tests/golden/make_spada_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/data/benchmarks/rotation_spada2011/polar_motion_test3-2.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spada2011_test3_2.json")


def main():
    rows, G = {}, {}
    for line in open(SRC):
        m = re.match(r"#\s+(cap|disc|point)\s*:\s*-\(([0-9.eE+-]+)\s*\+\s*i\s*([0-9.eE+-]+)\)", line)
        if m:
            G[m.group(1)] = [-float(m.group(2)), -float(m.group(3))]
            continue
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        load, cw, t = f[0], f[1], float(f[2])
        rows.setdefault(load, {}).setdefault(cw, []).append(
            {"t_kyr": t, "m_Vb_deg": float(f[3]), "m_Gs_deg": float(f[4])})
    json.dump({"source": "reference data/benchmarks/rotation_spada2011/polar_motion_test3-2.txt "
                         "(Spada et al. 2011, GJI 185, Table 14)",
               "G_sigma": G, "table": rows}, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
