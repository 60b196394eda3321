"""The oracle's six face fluxes against golden vectors from an independent
numpy restatement of the reference formulas (tests/golden/make_flux_golden.py):
the subsonic split polynomials, wave-speed estimates, star states and the
HLLCP/lowMach corrections that consistency and symmetry leave open."""
import os

import numpy as np
import pytest

SCHEMES = {"AUSM+": 0, "AUSM+up": 1, "HLL": 2, "HLLC": 3, "HLLCP": 4, "Kurganov": 5}
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "flux_faces.npz")


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_face_flux_golden(oracle, scheme):
    g = np.load(GOLDEN)
    for i in range(len(g["L"])):
        lm = int(g["low_mach"][i]) if scheme == "AUSM+up" else 0
        out = oracle.face_flux(SCHEMES[scheme], g["L"][i], g["R"][i], g["Sf"][i], gamma=float(g["gamma"]),
                               f_hllcp=float(g["f"][i]), low_mach=lm)
        ref = g[scheme][i]
        scale = np.abs(ref[:5]).max() + 1.0
        assert np.allclose(out[:5], ref[:5], rtol=0, atol=1e-11 * scale), (scheme, i, out[:5], ref[:5])
        uscale = np.abs(ref[5:]).max() + 1.0
        assert np.allclose(out[5:], ref[5:], rtol=0, atol=1e-11 * uscale), (scheme, i, out[5:], ref[5:])
