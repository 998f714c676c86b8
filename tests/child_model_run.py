"""Child process of the multi-process GPU test: runs `calls` coupling calls of
benchmark configuration `idx` at `lmax` (calls 40.. of its ice history) on
cuda:0 through the library and saves rsl, memory and the ocean mask.
Usage: python tests/child_model_run.py idx lmax calls out.npz"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main(idx, lmax, calls, out):
    import numpy as np
    from fastearth3d_amd import configs, lib
    c = configs.config(idx, lmax=lmax)
    se = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        se.enable_visc_3d(c["pert"])
    for k in range(calls):
        se.update(c["ice"](k + 40), c["dt_yr"])
    np.savez(out, rsl=se.rsl, memory=se.memory(), C=se.C)


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]), sys.argv[4])
