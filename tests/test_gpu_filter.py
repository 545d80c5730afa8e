"""filter_linelist on the device (korg_filter_linelist_batched, src/synthesize.jl:331-350) for
stars drawn from one master list in HBM.  Needs a GPU.

* 64 stars with 0-3 ranges each over a 300,000-line master (4000-7000 A): buffers of 0-15 A,
  ranges beyond the master, single-wavelength ranges, ranges whose buffers overlap.  Every star's
  lines equal the host filter_linelist's field by field (bit-exact; the host filter is checked
  against the reference's naive filter on the GALAH list, test/synthesize.jl:89-110) and select
  the naive filter's lines here too.  A star whose lines are one run of the master points into the
  master (no copy); the others' lines are gathered into a pool sized by a query call.
* The line opacity of 8 stars batched over the device-filtered lists equals, bitwise, the batch
  over the host-filtered lists uploaded, Float64 and Float32.
"""
import ctypes

import numpy as np
import pytest

from korg_line import _lib, workloads
from korg_line.atmosphere import Atmosphere
from korg_line.line_absorption import DeviceLineList, DeviceProblem
from korg_line.linelist import filter_linelist
from korg_line.wavelengths import Wavelengths

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def master():
    return workloads.synthetic_linelist(11, 300_000, 4000.0, 7000.0, (-4.0, 0.5), (0.0, 5.0))


def _stars(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nr = int(rng.integers(1, 4))
        lo = float(np.round(rng.uniform(3900.0, 7000.0), 2))
        spec = []
        for _ in range(nr):
            w = float(np.round(rng.uniform(0.0, 60.0), 2))
            spec.append((lo, lo + w))
            lo = float(np.round(lo + w + rng.uniform(0.01, 40.0), 2))
        out.append((Wavelengths(spec), float(rng.uniform(0.0, 15.0)) * 1e-8))
    # hand-made cases: beyond the master, one wavelength, buffers that overlap, a buffer of 0
    out[0] = (Wavelengths((7100.0, 7200.0)), 2e-8)
    out[1] = (Wavelengths((5000.0, 5000.0)), 21e-8)
    out[2] = (Wavelengths([(5000.0, 5001.0), (5002.0, 5003.0)]), 2e-8)
    out[3] = (Wavelengths([(4500.0, 4510.0), (6000.0, 6001.0)]), 0.0)
    return out


def _naive(ll, wls, b):
    keep = np.zeros(len(ll), dtype=bool)
    for lo, hi in wls.eachwindow():
        keep |= (lo - b <= ll.wl) & (ll.wl <= hi + b)
    return np.nonzero(keep)[0]


def _inside(ptr, base, n):
    return base <= ptr < base + 8 * n


def test_filter_matches_host_and_naive(ctx, master):
    stars = _stars(5, 64)
    dm = DeviceLineList(master)
    with pytest.warns(UserWarning):  # star 0 keeps nothing
        got = ctx.filter_linelist_batched(dm, [w for w, _ in stars], [b for _, b in stars])
    n_views = 0
    for k, (w, b) in enumerate(stars):
        ref = filter_linelist(master, w, b)
        idx = _naive(master, w, b)
        assert len(ref) == idx.size and np.array_equal(ref.wl, master.wl[idx]), k
        dev = got.to_host(k)
        assert got.n_lines(k) == len(ref), k
        for f in ("wl", "log_gf", "E_lower", "gamma_rad", "gamma_stark", "vdW1", "vdW2",
                  "species_index"):
            assert np.array_equal(getattr(dev, f), getattr(ref, f)), (k, f)
        if len(ref) and np.all(np.diff(idx) == 1):  # one run of the master: a view
            assert _inside(got.lines[k].wl, dm.t["wl"].ptr, len(master)), k
            n_views += 1
    assert got.n_lines(0) == 0 and got.n_lines(2) > 0
    assert n_views >= 30 and got.pool is not None
    # the pool size query reports the gathered lines exactly
    need = ctypes.c_int64(-1)
    fs = (_lib.KorgFilterStar * 1)()
    r = np.array(stars[3][0].eachwindow(), dtype=np.float64).ravel()
    fs[0] = _lib.KorgFilterStar(2, r.ctypes.data, 0.0)
    out = (_lib.KorgLines * 1)()
    rc = ctx._L.korg_filter_linelist_batched(ctx._h, ctypes.byref(dm.lines), 1, fs, None, 0, out,
                                             ctypes.byref(need))
    assert rc == _lib.KORG_ERR_CAPACITY and need.value == len(filter_linelist(master, stars[3][0],
                                                                               0.0))


def test_filter_rejects_bad_ranges(ctx, master):
    dm = DeviceLineList(master.take(np.arange(1000)))
    fs = (_lib.KorgFilterStar * 1)()
    r = np.array([5000e-8, 5010e-8, 5005e-8, 5020e-8])  # overlapping ranges
    fs[0] = _lib.KorgFilterStar(2, r.ctypes.data, 1e-8)
    out = (_lib.KorgLines * 1)()
    need = ctypes.c_int64(0)
    rc = ctx._L.korg_filter_linelist_batched(ctx._h, ctypes.byref(dm.lines), 1, fs, None, 0, out,
                                             ctypes.byref(need))
    assert rc == _lib.KORG_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("precision,dtype", [("fp64", "float64"), ("fp32", "float32")])
def test_batched_opacity_from_filtered_lists(ctx, master, precision, dtype):
    sun = Atmosphere.from_model("sun")
    rng = np.random.default_rng(8)
    stars = []
    for k in range(8):
        lo = float(np.round(rng.uniform(4000.0, 6800.0), 2))
        spec = [(lo, lo + 30.0)] if k % 2 == 0 else [(lo, lo + 15.0), (lo + 20.0, lo + 40.0)]
        atm = Atmosphere.scaled(float(rng.uniform(4500, 6500)), float(rng.uniform(2, 5)),
                                float(rng.uniform(-1.5, 0.3))) if k else sun
        stars.append((Wavelengths(spec), atm))
    dm = DeviceLineList(master)
    got = ctx.filter_linelist_batched(dm, [w for w, _ in stars], 10e-8)
    dev, host = [], []
    for k, (w, atm) in enumerate(stars):
        sub = filter_linelist(master, w, 10e-8)
        # the problem over the master's species table: the filtered lists index it
        dev.append(got.problem(k, workloads.make_problem(atm, master, w)))
        host.append(DeviceProblem(workloads.make_problem(atm, sub, w)))
        assert np.array_equal(dev[-1].problem.n_div_U, host[-1].problem.n_div_U)
    a_dev = [d.new_alpha(dtype) for d in dev]
    a_host = [d.new_alpha(dtype) for d in host]
    ctx.line_absorption_batched(a_dev, dev, precision)
    ctx.line_absorption_batched(a_host, host, precision)
    for k in range(len(stars)):
        x, y = a_dev[k].numpy(), a_host[k].numpy()
        assert np.max(y) > 0
        assert np.array_equal(x, y), k
