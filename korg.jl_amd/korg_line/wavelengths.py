"""Wavelength grids: a restatement of Korg's ``Wavelengths`` (``src/wavelengths.jl``).

A grid is a sorted list of evenly spaced ranges, given in Angstrom (values >= 1) or cm (values
< 1), and held in cm.  ``all_wls`` materialises every grid point the way Julia's twice-precision
``StepRangeLen`` does for ``range(start=a, stop=b, step=h) .* 1e-8``: the exact rational value of
each grid point (times the double nearest 1e-8 for Angstrom input), rounded once to the nearest
double.  The drop-in C ABI receives the caller's own ``all_wls`` array, so the device never
re-derives grid values; this class exists for the Python host mirror and the tests.
"""
from fractions import Fraction
import math

import numpy as np

_CM_PER_ANGSTROM = 1e-8


def _rat(x):
    """Julia's ``Base.rat``: a small rational n/d with Float64(n/d) == x (twiceprecision.jl)."""
    y = x
    a, b, c, d = 1, 0, 0, 1
    m = 16777216.0  # maxintfloat(Float32)
    while abs(y) <= m:
        f = int(y)  # trunc
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if b != 0 and float(a) / float(b) == x:
            break
        if y == 0:
            break
        y = 1.0 / y
    return a, b


def _range_values(start, stop, step):
    """Exact rational values (numerators over a common denominator) of ``start:step:stop``."""
    step_n, step_d = _rat(step)
    if step_d != 0 and step_n / step_d == step:
        start_n, start_d = _rat(start)
        stop_n, stop_d = _rat(stop)
        if start_d != 0 and stop_d != 0 and start_n / start_d == start and \
                stop_n / stop_d == stop:
            den = start_d * step_d // math.gcd(start_d, step_d)
            if den % start_d == 0 and den % step_d == 0:
                s_n = round(start * den)
                h_n = round(step * den)
                length = max(0, (den * stop_n - stop_d * s_n + h_n * stop_d) // (h_n * stop_d))
                if length >= 1:
                    return s_n, h_n, den, length
    # fall back to the literal start and step (exact binary values)
    fs, fh = Fraction(start), Fraction(step)
    length = int(math.floor((stop - start) / step + 1e-10)) + 1
    den = fs.denominator * fh.denominator // math.gcd(fs.denominator, fh.denominator)
    return int(fs * den), int(fh * den), den, length


def _materialise(s_n, h_n, den, length, scale):
    """Round (s_n + k h_n)/den * scale to the nearest double for k = 0..length-1."""
    sf = Fraction(scale)
    num_mul, den_mul = sf.numerator, sf.denominator * den
    out = np.empty(length, dtype=np.float64)
    for k in range(length):
        out[k] = ((s_n + k * h_n) * num_mul) / den_mul  # int/int true division: correctly rounded
    return out


class Wavelengths:
    """Possibly non-contiguous wavelength grid (``src/wavelengths.jl:29-131``).

    ``Wavelengths((5000, 5100))`` (0.01 A default step), ``Wavelengths((5000, 5100, 0.02))``,
    ``Wavelengths([(5000, 5010), (6000, 6010, 0.03)])``, or a single wavelength
    ``Wavelengths([5000.0])``.
    """

    def __init__(self, spec):
        if isinstance(spec, tuple):
            spec = [spec]
        spec = list(spec)
        if len(spec) == 0:
            raise ValueError("wavelengths must be non-empty")
        if not isinstance(spec[0], tuple):  # vector of wavelength values
            vals = [float(v) for v in spec]
            if len(vals) == 1:
                spec = [(vals[0], vals[0], 1.0)]
            else:
                d = np.diff(vals)
                if d.max() - d.min() > 1e-6:
                    raise ValueError("wavelengths are not linearly spaced to within 1e-06.")
                spec = [(vals[0], vals[-1], (vals[-1] - vals[0]) / (len(vals) - 1))]
        ranges = []
        for t in spec:
            if not 2 <= len(t) <= 3:
                raise ValueError(f"Each wavelength range must be (start, stop[, step]); got {t}")
            start, stop = float(t[0]), float(t[1])
            step = float(t[2]) if len(t) == 3 else (1e-10 if start < 1 else 0.01)
            ranges.append((start, stop, step))
        in_angstrom = ranges[0][0] >= 1
        scale = _CM_PER_ANGSTROM if in_angstrom else 1.0
        parts, self.wl_ranges = [], []
        for start, stop, step in ranges:
            s_n, h_n, den, length = _range_values(start, stop, step)
            vals = _materialise(s_n, h_n, den, length, scale)
            parts.append(vals)
            self.wl_ranges.append((float(vals[0]), step * scale, length))
        self.all_wls = np.concatenate(parts)
        if np.any(np.diff(self.all_wls) < 0):
            raise ValueError("wl_ranges must be sorted and non-overlapping")
        self.wl_ranges_angstrom = [(a if in_angstrom else a * 1e8, b if in_angstrom else b * 1e8)
                                   for a, b, _ in ranges]

    def __len__(self):
        return self.all_wls.size

    def __getitem__(self, i):
        return self.all_wls[i]

    def searchsortedfirst(self, lam):
        """1-based index of the first grid point >= lam (``src/wavelengths.jl:188-198``).

        As in the reference, a value >= 1 is read as Angstrom.
        """
        lam = lam * 1e-8 if lam >= 1 else lam
        return int(np.searchsorted(self.all_wls, lam, side="left")) + 1

    def searchsortedlast(self, lam):
        """1-based index of the last grid point <= lam, 0 if none (``src/wavelengths.jl:200-210``)."""
        lam = lam * 1e-8 if lam >= 1 else lam
        return int(np.searchsorted(self.all_wls, lam, side="right"))

    def eachwindow(self):
        """(first, last) wavelength of each range, cm (``src/wavelengths.jl:159``)."""
        out, i = [], 0
        for _, _, n in self.wl_ranges:
            out.append((float(self.all_wls[i]), float(self.all_wls[i + n - 1])))
            i += n
        return out

    def subspectrum_indices(self):
        out, i = [], 1
        for _, _, n in self.wl_ranges:
            out.append(range(i, i + n))
            i += n
        return out
