"""Pins of the oracle's limited reconstruction (reconstruct(rho|T) OSPRE,
reconstruct(U) OSPREV; foam LimitedScheme with the NVDTVD / NVDVTVDV
r-factor): an independent numpy restatement on a jittered mesh, the
second-order exactness for a linear field on a uniform grid (r = 1, OSPRE(1)
= 1: both sides give the linear interpolate), the closed-form OSPRE values,
and the reduction of OSPREV to OSPRE for a field along one direction."""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from test_oracle_sa import Geo


def ospre(r):
    rr = r * (r + 1.0)
    return 1.5 * rr / (rr + 1.0)


def rfac(gcf, gf):
    big = np.abs(gcf) >= 1000.0 * np.abs(gf)
    sgn = lambda x: np.where(x >= 0, 1.0, -1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(big, 2000.0 * sgn(gcf) * sgn(gf) - 1.0, 2.0 * gcf / gf - 1.0)
    return r


def numpy_recon(geo, cell, bnd):
    g = geo.grad(cell, bnd)
    nI = geo.nI
    P, N = cell[geo.own], cell[geo.nei]
    w = geo.g["weights"][:nI]
    d = geo.g["C"][geo.nei] - geo.g["C"][geo.own]
    if cell.ndim == 1:
        gf = N - P
        lo = ospre(rfac((d * g[geo.own]).sum(1), gf))
        ln = ospre(rfac((d * g[geo.nei]).sum(1), gf))
    else:
        gfv = N - P
        gf = (gfv * gfv).sum(1)
        dgo = np.einsum("fi,fij->fj", d, g[geo.own])
        dgn = np.einsum("fi,fij->fj", d, g[geo.nei])
        lo = ospre(rfac((gfv * dgo).sum(1), gf))
        ln = ospre(rfac((gfv * dgn).sum(1), gf))
    wo = lo * w + (1 - lo)
    wn = ln * w
    if cell.ndim == 2:
        wo, wn = wo[:, None], wn[:, None]
    own = np.concatenate([wo * (P - N) + N, bnd])
    nei = np.concatenate([wn * (P - N) + N, bnd])
    return own, nei


def test_ospre_closed_forms():
    assert ospre(1.0) == 1.0
    assert ospre(0.0) == 0.0 and ospre(-1.0) == 0.0
    assert ospre(2.0) == pytest.approx(1.5 * 6 / 7, rel=1e-15)
    assert ospre(1e6) == pytest.approx(1.5, rel=1e-5)


def test_recon_against_numpy(oracle):
    spec = cases.box_farfield_case(nx=20, ny=14, jitter=0.25)
    oc = oracle.OracleCase(spec)
    geo = Geo(oracle, spec.mesh)
    rng = np.random.default_rng(8)
    s = 1.0 + 0.2 * rng.uniform(-1, 1, geo.n)
    sb = 1.0 + 0.2 * rng.uniform(-1, 1, geo.nB)
    o1, n1 = oc.reconstruct(s, sb)
    o2, n2 = numpy_recon(geo, s, sb)
    assert np.abs(o1 - o2).max() < 1e-13 and np.abs(n1 - n2).max() < 1e-13
    v = rng.normal(size=(geo.n, 3))
    vb = rng.normal(size=(geo.nB, 3))
    o1, n1 = oc.reconstruct(v, vb)
    o2, n2 = numpy_recon(geo, v, vb)
    assert np.abs(o1 - o2).max() < 1e-12 and np.abs(n1 - n2).max() < 1e-12


def test_linear_field_is_reconstructed_exactly(oracle):
    spec = cases.box_farfield_case(nx=12, ny=10, jitter=0.0)
    oc = oracle.OracleCase(spec)
    geo = Geo(oracle, spec.mesh)
    C = geo.g["C"]
    Cf = geo.g["Cf"]
    a = np.array([3.0, -2.0, 0.0])
    s = 1.0 + C @ a
    sb = 1.0 + Cf[geo.nI:] @ a
    own, nei = oc.reconstruct(s, sb)
    exact = 1.0 + Cf[:geo.nI] @ a
    assert np.abs(own[:geo.nI] - exact).max() < 1e-13
    assert np.abs(nei[:geo.nI] - exact).max() < 1e-13
    # a vector along one direction: OSPREV gives the scalar OSPRE result
    rng = np.random.default_rng(9)
    t = rng.uniform(0.5, 1.5, geo.n)
    tb = rng.uniform(0.5, 1.5, geo.nB)
    e = np.array([0.6, 0.8, 0.0])
    ov, nv = oc.reconstruct(t[:, None] * e, tb[:, None] * e)
    os_, ns_ = oc.reconstruct(t, tb)
    assert np.abs(ov - os_[:, None] * e).max() < 1e-13
    assert np.abs(nv - ns_[:, None] * e).max() < 1e-13
