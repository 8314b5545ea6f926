"""ptg_histogram_noalloc through the C ABI against ptg_histogram on the same
resident final state: two calls in a row, the second over a smaller population
and fewer bins (the shared device array must be cleared between calls), count
histograms bit-exact, energy-weighted ones to 1e-12 of the largest bin (the
atomic summation order differs between launches), the computed range identical,
and the bin limit refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(eng, spec_args):
    from ptarmigan_amd.engine import Engine, _dptr
    out = []
    for fn in (eng._lib.ptg_histogram, eng._lib.ptg_histogram_noalloc):
        spec = Engine._hist_spec(*spec_args)
        bins = np.full(spec.nbins, np.nan)
        tot = C.c_double(np.nan)
        rc = fn(eng._ctx, C.byref(spec), _dptr(bins), C.byref(tot))
        assert rc == 0, eng._lib.ptg_last_error(eng._ctx)
        out.append((bins, spec.min, spec.max, tot.value))
    return out


def test_noalloc_histogram_matches_ptg_histogram():
    from ptarmigan_amd import Engine, Species, workloads
    from ptarmigan_amd.engine import _dptr
    PTG_Q_ENERGY, PTG_Q_ANGLE_X, W_NUMBER, W_ENERGY = 14, 0, 0, 1
    with Engine(0) as eng:
        laser, beam, opts = workloads.focused_lcfa_cascade(20_000)
        beam.radius = 1.5e-6
        eng.set_laser(laser)
        for call, (n, nbins) in enumerate(((20_000, 8192), (3_000, 150))):
            beam.n = n
            eng.build_beam(beam, seed=30 + call)
            cnt = eng.run(opts, seed=30 + call)
            say("noalloc call", call, "primaries", n, "bins", nbins, "final state", cnt["n_secondaries"])
            for sp, q, w in ((Species.PHOTON, PTG_Q_ENERGY, W_NUMBER), (Species.ELECTRON, PTG_Q_ENERGY, W_ENERGY),
                             (Species.PHOTON, PTG_Q_ANGLE_X, W_ENERGY)):
                (b1, lo1, hi1, t1), (b2, lo2, hi2, t2) = _both(eng, (sp, q, w, False, nbins))
                assert lo1 == lo2 and hi1 == hi2
                if w == W_NUMBER:
                    assert np.array_equal(b1, b2) and t1 == t2
                else:
                    assert np.allclose(b1, b2, rtol=0, atol=1e-12 * np.abs(b1).max())
                    assert t2 == pytest.approx(t1, rel=1e-12)
                assert b1.sum() > 0
        # beyond the fixed array: refused
        spec = Engine._hist_spec(Species.PHOTON, PTG_Q_ENERGY, W_NUMBER, False, 8193)
        bins = np.zeros(8193)
        tot = C.c_double(0.0)
        rc = eng._lib.ptg_histogram_noalloc(eng._ctx, C.byref(spec), _dptr(bins), C.byref(tot))
        assert rc == -1 and b"PTG_NOALLOC_MAX_BINS" in eng._lib.ptg_last_error(eng._ctx)
