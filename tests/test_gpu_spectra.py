"""Spectrum post-processing on the GPU (qs_broaden through the C ABI and
Spectra.Broaden of the API mirror) against the oracle's restatement of
BroadenG/L and BroadenGvar/Lvar: max |GPU - oracle| <= 1e-12 max |oracle|
(the sums over source points are split into chunks on the GPU, so the
rounding differs; the device exp differs from glibc's by at most an ulp)."""
import numpy as np
import pytest

import oracle
from quanty_amd.api import Spectra
from quanty_amd.engine import broaden, cf_spectrum

pytestmark = pytest.mark.gpu

TOL = 1e-12


def spectra(nspec, npts, seed, emin=-10.0, emax=20.0, gamma=0.4):
    """Continued-fraction spectra of random tridiagonal tables (a realistic input)."""
    rng = np.random.default_rng(seed)
    n = 40
    a = np.zeros((nspec, n + 1))
    a[:, 1:] = rng.uniform(-5, 15, (nspec, n))
    b = rng.uniform(0.2, 2.0, (nspec, n))
    return cf_spectrum(a, b, 0.0, emin, emax, npts - 1, gamma)


def close(got, ref):
    return np.max(np.abs(got - ref)) <= TOL * np.max(np.abs(ref))


@pytest.mark.parametrize("kind", ["G", "L"])
@pytest.mark.parametrize("per_point", [False, True])
@pytest.mark.parametrize("nspec", [1, 3, 11])
def test_broaden_matches_oracle(kind, per_point, nspec, ctx):
    npts = 2001
    dE = 30.0 / (npts - 1)
    G = spectra(nspec, npts, seed=nspec)
    rng = np.random.default_rng(7)
    f = rng.uniform(0.0, 1.2, npts) if per_point else 0.45
    got = broaden(G, dE, kind, f, ctx)
    ref = oracle.broaden(G, dE, kind, f, per_point=per_point)
    assert close(got, ref)


@pytest.mark.parametrize("npts", [1, 2, 7, 65, 130])
def test_broaden_small_grids(npts, ctx):
    rng = np.random.default_rng(npts)
    G = rng.standard_normal((9, npts)) + 1j * rng.standard_normal((9, npts))
    dE = 0.05
    for kind in ("G", "L"):
        assert close(broaden(G, dE, kind, 0.3, ctx), oracle.broaden(G, dE, kind, 0.3))
        if npts > 1:
            f = rng.uniform(0, 0.5, npts)
            assert close(broaden(G, dE, kind, f, ctx), oracle.broaden(G, dE, kind, f, per_point=True))


def test_broaden_narrower_than_grid_is_identity(ctx):
    G = spectra(2, 301, seed=4)
    dE = 30.0 / 300
    for kind in ("G", "L"):
        assert np.array_equal(broaden(G, dE, kind, 0.049, ctx), G)


def test_spectra_broaden_with_width_tables(ctx):
    """Spectra.Broaden(G, L): Gaussian, then Lorentzian, widths from energy tables
    (the form of the reference's XAS scripts: {{E_1, FWHM_1}, ...})."""
    npts, emin, emax = 3501, -13.5, 21.5
    G = spectra(3, npts, seed=11, emin=emin, emax=emax, gamma=0.07)
    S = Spectra(emin, emax, npts - 1, 0.07, G.copy())
    gt = [[-13.5, 0.46], [3, 0.46], [3.01, 0.8], [21.5, 0.9]]
    lt = [[-5, 0.2], [0, 0.3], [10, 0.6]]
    out = S.Broaden(gt, lt)
    assert out is S
    dE = (emax - emin) / (npts - 1)
    ref = oracle.broaden(G, dE, "G", oracle.fwhm_table(gt, emin, dE, npts), per_point=True)
    ref = oracle.broaden(ref, dE, "L", oracle.fwhm_table(lt, emin, dE, npts), per_point=True)
    assert close(S.G, ref)
    # numbers: one width each
    S2 = Spectra(emin, emax, npts - 1, 0.07, G.copy()).Broaden(0.5, 0.1)
    ref2 = oracle.broaden(oracle.broaden(G, dE, "G", 0.5), dE, "L", 0.1)
    assert close(S2.G, ref2)
