"""The coupling step on the Martinec et al. (2018) sea-level benchmark
(tests/test_benchmark_sle.f90; profiles data/benchmarks/sle_martinec2018,
committed as tests/golden/martinec2018.npz):

* at lmax 128, the four cases C2 (Heaviside cap on a shallow far basin,
  fixed ocean), D3 (growing cap, deep near basin, fixed ocean), E2 (as D3
  with a migrating coastline) and F1 (as E2 with the paleotopography
  spin-up) through the HIP library, against the published u, geoid, sea
  surface and sea level profiles with the reference's tolerances
  (peak-normalised max error 30 %, 12 %, 8 %, 12 %);
* at lmax 32, the whole history of C2 (500 calls) and E2 (750 calls)
  through the HIP library against the oracle: ocean function bit for bit,
  rsl and the u / N spectra within 1e-8, which extends the full-run parity
  to the fixed-ocean sea-level solve, the growing cap and the basin
  topography.
"""
import numpy as np
import pytest

import martinec as M
from conftest import rel
from fastearth3d_amd import _abi, configs

pytestmark = pytest.mark.gpu


def _gpu(params, z, h):
    from fastearth3d_amd import lib
    return lib.SolidEarth(params, z, h)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("case", ["C2", "D3", "E2", "F1"])
def test_martinec_published_profiles(case):
    spectra, passes = M.run_case(_gpu, case, 128, log=print)
    errs = M.compare(case, spectra, 128)
    for fig, e in errs.items():
        print(case, fig, " ".join("%s %.2f%%" % (k, 100 * v) for k, v in e.items()), flush=True)
    if case == "F1":
        assert passes < M.SPIN_MAX
    for fig, e in errs.items():
        for q, v in e.items():
            assert v <= M.TOL[q], (case, fig, q, v)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", ["C2", "E2"])
def test_martinec_history_vs_oracle(oracle, case):
    c = configs.martinec2018(case, 32)
    g = _gpu(c["params"], c["basin"], c["h_ice_eq"])
    o = oracle.OracleModel(c["params"], c["basin"], c["h_ice_eq"])
    for k in range(c["steps"]):
        h = c["ice"](k)
        g.update(h, c["dt_yr"])
        o.update(h, c["dt_yr"])
        if k % 50 == 49 or k == c["steps"] - 1:
            assert np.array_equal(g.C, o.C), (case, k)
            assert rel(g.rsl, o.rsl) <= 1e-8, (case, k, rel(g.rsl, o.rsl))
    for sp in (_abi.SP_U, _abi.SP_N, _abi.SP_SIGMA):
        assert rel(g.spectrum(sp), o.spectrum(sp)) <= 1e-8, (case, sp)
    for m in (g, o):
        m.finalize()
