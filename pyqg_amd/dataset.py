"""A model's state, constants and time-averaged diagnostics as labelled arrays:
the reference's xarray output (pyqg/xarray_output.py:1-174, Model.to_dataset at
pyqg/model.py:840-848).

`dataset_parts(m)` assembles the variables, coordinates and global attributes
as plain ``(dims, data, attrs)`` tuples from the model's host views, so the
assembly needs nothing beyond numpy.  `model_to_dataset(m)` hands them to
``xarray.Dataset``; xarray is an optional dependency, and without it
`model_to_dataset` raises ImportError as the reference's module does on import.

The variable names, dimension names, units and long names are the output
format that readers of the reference's datasets rely on, so they are kept as
the reference writes them (including its spelling)."""
import numpy as np

from .errors import DiagnosticNotFilledError

_SPATIAL = ('time', 'lev', 'y', 'x')
_SPECTRAL = ('time', 'lev', 'l', 'k')

# name -> (dims, units, long name), in the reference's order
# (xarray_output.py:14-47)
_VARIABLES = {
    'q': (_SPATIAL, 's^-1', 'potential vorticity in real space'),
    'u': (_SPATIAL, 'm s^-1', 'zonal velocity anomaly'),
    'v': (_SPATIAL, 'm s^-1', 'meridional velocity anomaly'),
    'ufull': (_SPATIAL, 'm s^-1', 'zonal full velocities in real space'),
    'vfull': (_SPATIAL, 'm s^-1', 'meridional full velocities in real space'),
    'qh': (_SPECTRAL, 's^-1', 'potential vorticity in spectral space'),
    'uh': (_SPECTRAL, 'm s^-1', 'zonal velocity anomaly in spectral space'),
    'vh': (_SPECTRAL, 'm s^-1', 'meridional velocity anomaly in spectral space'),
    'ph': (_SPECTRAL, 'm^2 s^-1', 'streamfunction in spectral space'),
    'dqhdt': (_SPECTRAL, 's^-2',
              'previous partial derivative of potential vorticity wrt. time in spectral space'),
    'Ubg': ('lev', 'm s^-1', 'background zonal velocity'),
    'Qy': ('lev', 'm^-1 s^-1', 'background potential vorticity gradient'),
}

# spectral variable -> (real-space name, units, long name) (xarray_output.py:49-53)
_INVERTED = {
    'dqhdt': ('dqdt', 's^-2',
              'previous partial derivative of potential vorticity wrt. time in real space'),
    'ph': ('p', 'm^2 s^-1', 'streamfunction in real space'),
}

# coordinate -> (attrs, values from the model) (xarray_output.py:56-116)
_COORDINATES = {
    'time': ({'long_name': 'model time', 'units': 's'}, lambda m: np.array([m.t])),
    'lev': ({'long_name': 'vertical levels'}, lambda m: np.arange(1, m.nz + 1)),
    'lev_mid': ({'long_name': 'vertical level interface'}, lambda m: np.arange(1.5, m.nz + .5)),
    'x': ({'long_name': 'real space grid points in the x direction', 'units': 'grid point'},
          lambda m: m.x[0, :]),
    'y': ({'long_name': 'real space grid points in the y direction', 'units': 'grid point'},
          lambda m: m.y[:, 0]),
    'l': ({'long_name': 'spectal space grid points in the l direction',
           'units': 'meridional wavenumber'}, lambda m: m.ll),
    'k': ({'long_name': 'spectal space grid points in the k direction',
           'units': 'zonal wavenumber'}, lambda m: m.kk),
}

# model attributes copied as global attributes "pyqg:<name>" when present
# (xarray_output.py:79-106)
_GLOBAL_ATTRS = ('beta', 'delta', 'del2', 'dt', 'filterfac', 'L', 'M', 'nk', 'nl', 'ntd', 'nx',
                 'ny', 'nz', 'pmodes', 'radii', 'rd', 'rho', 'rek', 'taveint', 'tavestart',
                 'tc', 'tmax', 'tsnapint', 'tsnapstart', 'twrite', 'W')

TITLE = 'pyqg: Python Quasigeostrophic Model'
REFERENCE = 'https://pyqg.readthedocs.io/en/latest/index.html'


def _with_time(dims, data):
    """A leading time axis of length one (snapshots are stacked along it)."""
    return (('time',) + tuple(dims), np.asarray(data)[np.newaxis, ...])


def dataset_parts(m):
    """(variables, coordinates, attrs) of `m`, each variable and coordinate a
    ``(dims, data, attrs)`` tuple (xarray_output.py:118-168).  Variables:
    the state and background fields, p and dqdt transformed back from ph and
    dqhdt, and every diagnostic that has been accumulated (with a time axis
    added where its dims have none)."""
    variables = {}
    for name, (dims, units, long_name) in _VARIABLES.items():
        if not hasattr(m, name):
            continue
        data = np.array(getattr(m, name))
        attrs = {'units': units, 'long_name': long_name}
        if 'time' in dims:
            variables[name] = (dims, data[np.newaxis, ...], attrs)
        else:
            variables[name] = (dims, data, attrs)
    for spectral, (name, units, long_name) in _INVERTED.items():
        if hasattr(m, spectral):
            data = np.asarray(m.ifft(getattr(m, spectral)))
            variables[name] = (_SPATIAL, data[np.newaxis, ...],
                               {'units': units, 'long_name': long_name})

    coordinates = {}
    for name, (attrs, values) in _COORDINATES.items():
        coordinates[name] = ((name,), np.array(values(m)), dict(attrs))

    attrs = {f'pyqg:{name}': getattr(m, name) for name in _GLOBAL_ATTRS if hasattr(m, name)}

    for name, entry in m.diagnostics.items():
        try:
            data = m.get_diagnostic(name)
        except DiagnosticNotFilledError:
            continue
        if np.ndim(data) == 0:
            data = np.array([data])
        dims = tuple(entry.get('dims') or ())
        dattrs = {'long_name': entry.get('description'), 'units': entry.get('units')}
        if 'time' in dims:
            variables[name] = (dims, np.asarray(data), dattrs)
        else:
            variables[name] = _with_time(dims, data) + (dattrs,)
    return variables, coordinates, attrs


def model_to_dataset(m):
    """``xarray.Dataset`` of `m` (xarray_output.py:118-174)."""
    try:
        import xarray as xr
    except ImportError as e:
        raise ImportError(
            "Xarray output in pyqg_amd requires the xarray package, which is not installed "
            "on this system. Install xarray to use Model.to_dataset().") from e
    variables, coordinates, attrs = dataset_parts(m)
    ds = xr.Dataset(variables, coords=coordinates, attrs=attrs)
    ds.attrs['title'] = TITLE
    ds.attrs['reference'] = REFERENCE
    return ds
