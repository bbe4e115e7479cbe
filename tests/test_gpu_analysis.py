"""On-device analysis against the CPU: the kinetic energy (kinetic
temperature), the pair-distance histogram behind g(r), and the trajectory
spectrum of the reference's FFTAnalysis -a (through demon_trajectory_spectrum
and the demon_fft tool on a DEMON output file).

Bars: kinetic energies within 1e-13 relative of numpy's; histogram counts
equal (integers, the same squared-distance bins); spectra within 1e-12 of the
series' absolute sum against the oracle's summed transforms.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import demon_simulation_code_amd as dm
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, charges, lattice

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
FFT = HERE.parent / "demon_simulation_code_amd" / "demon_fft"
HOST = HERE / "host" / "test_host"


def _ctx(x, y, vx, vy, m):
    n = len(x)
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA)
    c.upload(x, y, vx, vy, charges(n), m)
    return c


def numpy_hist(x, y, r_max, nbins):
    i, j = np.triu_indices(len(x), 1)
    dx, dy = x[i] - x[j], y[i] - y[j]
    d2 = dx * dx + dy * dy
    w = r_max / nbins
    e = np.arange(nbins + 1) * w
    e[-1] = r_max
    e2 = e * e
    d2 = d2[d2 < e2[-1]]
    b = np.searchsorted(e2, d2, side="right") - 1
    return np.bincount(b, minlength=nbins).astype(np.uint64)


@pytest.mark.parametrize("n", [100, 1024, 3000])
def test_kinetic_energy(n):
    rng = np.random.default_rng(n)
    x, y = lattice(n, seed=2)
    vx, vy = rng.normal(0.01, 1e-3, (2, n))
    m = M0 * rng.uniform(0.5, 1.5, n)
    with _ctx(x, y, vx, vy, m) as c:
        total, thermal = c.kinetic_energy()
    want = 0.5 * np.sum(m * (vx * vx + vy * vy))
    vcx, vcy = np.sum(m * vx) / np.sum(m), np.sum(m * vy) / np.sum(m)
    want_th = 0.5 * np.sum(m * ((vx - vcx) ** 2 + (vy - vcy) ** 2))
    assert total == pytest.approx(want, rel=1e-13)
    assert thermal == pytest.approx(want_th, rel=1e-12)
    assert thermal < 0.05 * total   # the drift of 0.01 m/s dominates


@pytest.mark.parametrize("n,r_max,nbins", [(100, 3e-3, 60), (1024, 2e-3, 1), (1000, 5e-3, 2048),
                                            (3000, 1.2e-3, 100), (777, 1.0, 37)])
def test_pair_histogram_exact(n, r_max, nbins):
    x, y = lattice(n, seed=n)
    z = np.zeros(n)
    with _ctx(x, y, z, z, np.full(n, M0)) as c:
        got = c.pair_histogram(r_max, nbins)
        again = c.pair_histogram(r_max, nbins)
    want = numpy_hist(x, y, r_max, nbins)
    assert np.array_equal(got, want)
    assert np.array_equal(got, again)
    if r_max == 1.0:
        assert int(got.sum()) == n * (n - 1) // 2


def test_pair_histogram_full_size():
    """N = 131072: every pair below a radius larger than the cloud, and the
    short-range histogram equal to one built from a k-d tree's pair list."""
    from scipy.spatial import cKDTree

    n, r_max, nbins = 131072, 1.2e-3, 60
    x, y = lattice(n, seed=3)
    z = np.zeros(n)
    with _ctx(x, y, z, z, np.full(n, M0)) as c:
        all_pairs = c.pair_histogram(1.0, 16)
        near = c.pair_histogram(r_max, nbins)
    assert int(all_pairs.sum()) == n * (n - 1) // 2
    pairs = cKDTree(np.c_[x, y]).query_pairs(r_max * (1 + 1e-9), output_type="ndarray")
    dx, dy = x[pairs[:, 0]] - x[pairs[:, 1]], y[pairs[:, 0]] - y[pairs[:, 1]]
    d2 = dx * dx + dy * dy
    e = np.arange(nbins + 1) * (r_max / nbins)
    e[-1] = r_max
    e2 = e * e
    d2 = d2[d2 < e2[-1]]
    want = np.bincount(np.searchsorted(e2, d2, side="right") - 1, minlength=nbins)
    assert np.array_equal(near, want.astype(np.uint64))
    assert int(near.sum()) > 8 * n


def test_trajectory_spectrum_vs_oracle(oracle):
    rng = np.random.default_rng(4)
    for rows, stride, nuse in ((500, 40, 40), (333, 17, 5), (1, 3, 3), (2048, 64, 64)):
        t = np.arange(rows)[:, None]
        series = 1e-3 * np.cos(2 * np.pi * 7 * t / max(rows, 1) + rng.uniform(0, 1, stride)) \
            + 1e-4 * rng.normal(size=(rows, stride))
        got = dm.trajectory_spectrum(series, nuse)
        want = oracle.average_fft(series, nuse)
        assert got.shape == want.shape == (rows // 2,)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(series[:, :nuse]).sum() / nuse)


def test_demon_fft_tool_on_a_demon_file(tmp_path, oracle):
    """demon_fft -a N -i file -o out: the reference's FFTAnalysis output format
    (frequency i / (rows 0.01), amplitude) from a DEMON file's X_POSITION rows."""
    path = tmp_path / "data.fits"
    subprocess.run([str(HOST), "fits_write", str(path), "40", "64", str(1 | 2 | 4)], check=True, timeout=120)
    out = tmp_path / "output.dat"
    r = subprocess.run([str(FFT), "-a", "25", "-i", str(path), "-o", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [ln.split("\t") for ln in out.read_text().splitlines()]
    assert len(rows) == 32
    # the file's rows: row k has x + 1e-6 k (test_host fits_write)
    lines = subprocess.run([str(HOST), "fits_read", str(path), "0"], capture_output=True, text=True,
                           check=True).stdout.splitlines()
    nk = int(lines[1])
    x0 = np.array([float(ln.split()[2]) for ln in lines[2 + nk:]])
    series = x0[None, :25] + 1e-6 * np.arange(64)[:, None]
    want = oracle.average_fft(series, 25)
    for i, (f, a) in enumerate(rows):
        assert f == "%g" % (i / (64 * 0.01))
        assert float(a) == pytest.approx(want[i], rel=1e-5, abs=1e-12 * np.abs(series).sum() / 25)
