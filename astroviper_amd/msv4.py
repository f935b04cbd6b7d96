"""A minimal MSv4-shaped container for environments without xarray/xradio.

The reference entry points take ``xarray.DataTree`` measurement sets
(xradio MSv4).  When xarray is importable those are accepted as they are;
otherwise ``VisibilityDataset`` offers the attributes the gridder touches:
``ds[name].data`` (a mutable NumPy array), ``ds.UVW``, ``ds.WEIGHT``,
``ds.FLAG``, ``ds.coords[...]``, ``ds.dims``, ``ds.sizes``, membership
tests and item assignment.  ``ImageDataset`` does the same for the
(time, frequency, polarization, l|u, m|v) image datasets of
``grid2xradio_spheroid_ms4`` and ``make_psf``.
"""
from __future__ import annotations

from typing import Dict, Iterable, Optional

import numpy as np

try:  # pragma: no cover - xarray is absent in the build container
    import xarray as _xr
except Exception:  # noqa: BLE001
    _xr = None


class Coord:
    def __init__(self, values, dims=None, attrs=None):
        self.values = np.asarray(values)
        self.dims = tuple(dims) if dims is not None else ()
        self.attrs = dict(attrs or {})

    @property
    def data(self):
        return self.values

    def __len__(self):
        return len(self.values)

    def __getitem__(self, k):
        return Coord(self.values[k])


class DataArray:
    """An ndarray with named dimensions and coordinates."""

    def __init__(self, data, coords=None, dims=None, name=None, attrs=None):
        self.data = np.asarray(data) if not isinstance(data, np.ndarray) else data
        if dims is None:
            dims = tuple(f"dim_{i}" for i in range(self.data.ndim))
        self.dims = tuple(dims)
        if len(self.dims) != self.data.ndim:
            raise ValueError(f"dims {self.dims} do not match shape {self.data.shape}")
        self.coords: Dict[str, Coord] = {}
        if coords:
            for k, v in dict(coords).items():
                self.coords[k] = v if isinstance(v, Coord) else Coord(v)
        self.name = name
        self.attrs = dict(attrs or {})

    @property
    def values(self):
        return self.data

    @property
    def shape(self):
        return self.data.shape

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def sizes(self):
        return dict(zip(self.dims, self.data.shape))

    def __array__(self, dtype=None, copy=None):
        return self.data if dtype is None else self.data.astype(dtype)

    def _wrap(self, data):
        return DataArray(data, coords=self.coords, dims=self.dims)

    def __sub__(self, o):
        return self._wrap(self.data - (o.data if isinstance(o, DataArray) else o))

    def __add__(self, o):
        return self._wrap(self.data + (o.data if isinstance(o, DataArray) else o))

    def __mul__(self, o):
        return self._wrap(self.data * (o.data if isinstance(o, DataArray) else o))

    def __imul__(self, o):
        self.data *= (o.data if isinstance(o, DataArray) else o)
        return self

    def __getitem__(self, k):
        return self.data[k]

    def __setitem__(self, k, v):
        self.data[k] = v

    def max(self):
        return self.data.max()

    def isel(self, **idx):
        sl = tuple(idx.get(d, slice(None)) for d in self.dims)
        keep = [d for d, s in zip(self.dims, sl) if not isinstance(s, (int, np.integer))]
        return DataArray(self.data[sl], dims=keep,
                         coords={k: v for k, v in self.coords.items() if k in keep})

    def sel(self, **labels):
        """Label selection on 1-D coordinates; scalar labels give views."""
        idx = {}
        for d, lab in labels.items():
            vals = self.coords[d].values
            hits = np.nonzero(vals == lab)[0]
            if hits.size == 0:
                raise KeyError(f"{lab!r} not in coordinate {d}")
            idx[d] = int(hits[0])
        return self.isel(**idx)


class _Dataset:
    def __init__(self, data_vars: Optional[Dict[str, DataArray]] = None,
                 coords: Optional[Dict[str, object]] = None, attrs=None):
        object.__setattr__(self, "data_vars", dict(data_vars or {}))
        object.__setattr__(self, "coords", {})
        for k, v in dict(coords or {}).items():
            self.coords[k] = v if isinstance(v, Coord) else Coord(v, dims=(k,))
        object.__setattr__(self, "attrs", dict(attrs or {}))

    def __getitem__(self, name):
        if name in self.data_vars:
            return self.data_vars[name]
        if name in self.coords:
            return self.coords[name]
        raise KeyError(name)

    def __setitem__(self, name, value):
        if not isinstance(value, DataArray):
            raise TypeError("assign a DataArray")
        self.data_vars[name] = value

    def __contains__(self, name):
        return name in self.data_vars

    def __getattr__(self, name):
        dv = object.__getattribute__(self, "data_vars")
        if name in dv:
            return dv[name]
        co = object.__getattribute__(self, "coords")
        if name in co:
            return co[name]
        raise AttributeError(name)

    @property
    def dims(self):
        out = {}
        for da in self.data_vars.values():
            out.update(da.sizes)
        for k, c in self.coords.items():
            if c.dims == (k,):
                out.setdefault(k, len(c.values))
        return out

    @property
    def sizes(self):
        return self.dims


class VisibilityDataset(_Dataset):
    """MSv4-like visibility dataset: VISIBILITY, WEIGHT, FLAG over
    (time, baseline_id, frequency, polarization) and UVW over
    (time, baseline_id, uvw_label)."""


class ImageDataset(_Dataset):
    """Image dataset over (time, frequency, polarization, l, m) or (u, v)."""


def is_datatree(obj) -> bool:
    """True for the measurement-set types the entry points accept."""
    if isinstance(obj, VisibilityDataset):
        return True
    if _xr is not None:  # pragma: no cover
        dt = getattr(getattr(_xr, "core", None), "datatree", None)
        if dt is not None and isinstance(obj, dt.DataTree):
            return True
        if hasattr(_xr, "DataTree") and isinstance(obj, _xr.DataTree):
            return True
    return False


def make_data_array(data, coords, dims):
    """DataArray of the installed flavour (xarray when present)."""
    if _xr is not None:  # pragma: no cover
        return _xr.DataArray(data, coords=coords, dims=dims)
    return DataArray(data, coords=coords, dims=dims)


def make_visibility_dataset(vis, uvw, weight, flag, freq, times=None, pols=None,
                            baselines=None) -> VisibilityDataset:
    """Wrap arrays shaped like an MSv4 partition."""
    nt, nb, nf, npol = weight.shape
    times = np.arange(nt, dtype=np.float64) if times is None else np.asarray(times)
    pols = (np.array(["XX", "XY", "YX", "YY"][:npol] if npol <= 4 else np.arange(npol))
            if pols is None else np.asarray(pols))
    baselines = np.arange(nb) if baselines is None else np.asarray(baselines)
    coords = {"time": times, "baseline_id": baselines, "frequency": np.asarray(freq),
              "polarization": pols, "uvw_label": np.array(["u", "v", "w"])}
    vdims = ("time", "baseline_id", "frequency", "polarization")
    dv = {
        "VISIBILITY": DataArray(vis, dims=vdims, coords={k: coords[k] for k in vdims}),
        "WEIGHT": DataArray(weight, dims=vdims, coords={k: coords[k] for k in vdims}),
        "FLAG": DataArray(flag, dims=vdims, coords={k: coords[k] for k in vdims}),
        "UVW": DataArray(uvw, dims=("time", "baseline_id", "uvw_label"),
                         coords={k: coords[k] for k in ("time", "baseline_id", "uvw_label")}),
    }
    return VisibilityDataset(dv, coords)


def datasets(vis) -> Iterable:
    return [vis] if is_datatree(vis) else list(vis)
