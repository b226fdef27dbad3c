"""Colour normalisations for plotting Wigner functions (wigner_normalize.py).

A Wigner function takes both signs, so the useful colour maps put W = 0 at the middle of
the map.  ``WignerNormalize`` maps [vmin, 0, vmax] piecewise linearly onto [0, 0.5, 1].
``WignerSymLogNorm`` first applies a symmetric logarithm (linear for |x| < linthresh) and
then maps the same way, so that small negative regions stay visible.  The reference's
``clip`` branch uses ``ma`` without importing it; here clipping works.
"""
import numpy as np
from matplotlib.colors import Normalize


def _piecewise(x, lo, hi):
    """[lo, 0, hi] -> [0, 0.5, 1], linear on each side (lo <= 0 <= hi assumed)."""
    return np.interp(x, (lo, 0., hi), (0., 0.5, 1.))


class WignerNormalize(Normalize):
    """Normalize with W = 0 mapped to 0.5 (wigner_normalize.py:5-45)."""

    def __call__(self, value, clip=None):
        if clip is None:
            clip = self.clip
        result, is_scalar = self.process_value(value)
        self.autoscale_None(result)
        vmin, vmax = float(self.vmin), float(self.vmax)
        if vmin > vmax:
            raise ValueError("minvalue must be less than or equal to maxvalue")
        if vmin == vmax:
            result.fill(0)
        else:
            data = np.ma.getdata(result)
            if clip:
                data = np.clip(data, vmin, vmax)
            result = np.ma.array(_piecewise(data, vmin, vmax), mask=np.ma.getmask(result), copy=False)
        if is_scalar:
            result = result[0]
        return result


class WignerSymLogNorm(Normalize):
    """Symmetric-log normalisation with W = 0 mapped to 0.5 (wigner_normalize.py:48-88).

    The transform is y = x / linthresh for |x| <= linthresh and
    y = sign(x) (1 + log10(|x| / linthresh)) beyond: continuous, odd and monotonic."""

    def __init__(self, linthresh, vmin=None, vmax=None, clip=False):
        super().__init__(vmin=vmin, vmax=vmax, clip=clip)
        if linthresh <= 0:
            raise ValueError("linthresh must be positive")
        self.linthresh = float(linthresh)

    def _transform(self, x):
        a = np.abs(x) / self.linthresh
        with np.errstate(divide="ignore"):
            y = np.where(a <= 1.0, a, 1.0 + np.log10(np.maximum(a, 1.0)))
        return np.sign(x) * y

    def __call__(self, value, clip=None):
        if clip is None:
            clip = self.clip
        result, is_scalar = self.process_value(value)
        self.autoscale_None(result)
        vmin, vmax = float(self.vmin), float(self.vmax)
        if vmin > vmax:
            raise ValueError("minvalue must be less than or equal to maxvalue")
        if vmin == vmax:
            result.fill(0)
        else:
            data = np.ma.getdata(result)
            if clip:
                data = np.clip(data, vmin, vmax)
            lo, hi = self._transform(np.array([vmin, vmax]))
            mapped = _piecewise(self._transform(data), min(lo, 0.), max(hi, 0.))
            result = np.ma.array(mapped, mask=np.ma.getmask(result), copy=False)
        if is_scalar:
            result = result[0]
        return result
