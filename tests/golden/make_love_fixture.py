"""Extract the M3-L70-V01 normal-mode Love-number table shipped with the
reference (data/benchmarks/love_M3-L70-V01/mod_M3-L70-V01, giapy / Spada et al.
2011 benchmark; see data/benchmarks/PROVENANCE.md there) into a small JSON
fixture: per degree the elastic (t=0) and fluid (t->inf) loading Love numbers
h, l, k, and for degrees up to 64 the normal modes: per mode the relaxation
rate s [1/kyr] and the residues k_i, h_i, l_i, so that under a Heaviside load
x(t) = x_e + sum_i x_i (1 - exp(s_i t)) (the fluid limit is x_e + sum_i x_i).
File format per the reference reader (tests/test_benchmark_love.f90:
read_ref): 5 earth-model lines, then per degree a header `n nmodes k h l`
(elastic), nmodes mode lines `n s k h l`, and a fluid line `n nmodes k h l`.

Run once in the build container: python tests/golden/make_love_fixture.py
"""
import json
import os
import sys

SRC = "/root/reference/data/benchmarks/love_M3-L70-V01/mod_M3-L70-V01"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "love_M3-L70-V01.json")


def main():
    lines = [l.split() for l in open(SRC) if l.strip()]
    rows = lines[5:]
    i, out = 0, {}
    while i < len(rows):
        n, nm = int(rows[i][0]), int(rows[i][1])
        ke, he, le = map(float, rows[i][2:5])
        fl = rows[i + 1 + nm]
        kf, hf, lf = map(float, fl[2:5])
        out[n] = {"elastic": {"h": he, "l": le, "k": ke}, "fluid": {"h": hf, "l": lf, "k": kf}}
        if n <= 64:
            out[n]["modes"] = [[float(v) for v in rows[i + 1 + q][1:5]] for q in range(nm)]
        i += 2 + nm
    json.dump({"source": "reference data/benchmarks/love_M3-L70-V01/mod_M3-L70-V01",
               "degrees": out}, open(OUT, "w"), indent=0)
    print("wrote", OUT, len(out), "degrees")


if __name__ == "__main__":
    sys.exit(main())
