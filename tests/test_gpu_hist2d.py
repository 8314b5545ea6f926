"""ptg_histogram_2d follows Histogram::generate_2d (src/output/hgram.rs:326-446):
each axis's range is its global min and max, a pair is binned only when both
values are finite, the flat bin is by * ny + bx (the reference's row-major
indexing with the y bin count; rows overlap when nx != ny) and indices past
nx * ny are dropped (the maximum of each axis falls off the end), and the total
counts every weight. Checked against numpy's bincount on the same particles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M_E = 0.510999


def _ref(x, y, nx, ny):
    bx = np.floor((x - x.min()) / ((x.max() - x.min()) / nx)).astype(np.int64)
    by = np.floor((y - y.min()) / ((y.max() - y.min()) / ny)).astype(np.int64)
    f = by * ny + bx
    return np.bincount(f[f < nx * ny], minlength=nx * ny).astype(float)


@pytest.mark.parametrize("nx,ny", [(200, 200), (50, 30), (17, 64)])
def test_histogram_2d_matches_numpy(engine, nx, ny):
    from ptarmigan_amd import Species, workloads
    print("hist2d", nx, ny, flush=True)
    laser, beam, opts = workloads.focused_lcfa_cascade(2000)
    beam.radius = 1.5e-6
    engine.set_laser(laser)
    engine.build_beam(beam, seed=9)
    cnt = engine.run(opts, seed=9)
    sec, _ = engine.fetch(cnt)
    ph = sec["species"] == 2
    e = M_E * sec["u"][ph, 0]
    px = M_E * sec["u"][ph, 1]
    counts, (xlo, xhi), (ylo, yhi), total = engine.histogram2d(Species.PHOTON, 14, 4, nx, ny)  # energy, p_x
    assert (xlo, xhi) == (e.min(), e.max()) and (ylo, yhi) == (px.min(), px.max())
    assert np.array_equal(counts, _ref(e, px, nx, ny))
    assert total == pytest.approx(sec["weight"][ph].sum())
