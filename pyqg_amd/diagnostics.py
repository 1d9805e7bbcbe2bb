"""Time-averaged diagnostics of the models, accumulated on the device.

pyqg keeps a registry of diagnostics (pyqg/model.py:630-822): every entry has
a function of the model, and between tavestart and the end of a run the
function is evaluated every taveint and summed on the host.  Here the
built-in entries (model.py:644-736, qg_diagnostics.py:6-70,
qg_model.py:226-263, layered_model.py:256-301) are summed by libqgk on the
GPU, per member, from the fields the step has just computed; the registry
only keeps the counts.  Entries added with add_diagnostic keep pyqg's host
evaluation.  The registry's dictionaries, add_diagnostic, describe_diagnostics
and get_diagnostic behave as pyqg's.
"""
import numpy as np

from . import constants as C
from .errors import DiagnosticNotFilledError

# name: (description, units, dims, extra registry keys)
CORE = {
    "Ensspec": ("enstrophy spectrum", "s^-2", ("lev", "l", "k"), {}),
    "KEspec": ("kinetic energy spectrum", "m^2 s^-2", ("lev", "l", "k"), {}),
    "EKEdiss": ("total energy dissipation by bottom drag", "m^2 s^-3", ("time",), {}),
    "KEfrictionspec": ("total energy dissipation spectrum by bottom drag", "m^2 s^-3",
                       ("l", "k"), {}),
    "EKE": ("mean eddy kinetic energy", "m^2 s^-2", ("lev",), {}),
    "Dissspec": ("Spectral contribution of filter dissipation to total energy", "m^2 s^-3",
                 ("l", "k"), {}),
    "ENSDissspec": ("Spectral contribution of filter dissipation to barotropic enstrophy",
                    "s^-3", ("l", "k"), {}),
    "paramspec": ("Spectral contribution of subgrid parameterization to energy (if present)",
                  "m^2 s^-3", ("l", "k"), {"skip_comparison": True}),
    "ENSparamspec": ("Spectral contribution of subgrid parameterization to enstrophy", "s^-3",
                     ("l", "k"), {"skip_comparison": True}),
}
QGDIAG = {
    "entspec": ("barotropic enstrophy spectrum", "m s^-2", ("l", "k"), {}),
    "paramspec_APEflux": ("total additional APE flux due to subgrid parameterization",
                          "m^2 s^-3", ("l", "k"), {"skip_comparison": True}),
    "paramspec_KEflux": ("total additional KE flux due to subgrid parameterization",
                         "m^2 s^-3", ("l", "k"), {"skip_comparison": True}),
    "ENSflux": ("barotropic enstrophy flux", "s^-3", ("l", "k"), {}),
    "ENSgenspec": ("the spectrum of the rate of generation of barotropic enstrophy", "s^-3",
                   ("l", "k"), {}),
    "ENSfrictionspec": ("the spectrum of the rate of dissipation of barotropic enstrophy due "
                        "to bottom friction", "s^-3", ("l", "k"), {}),
    "APEgenspec": ("the spectrum of the rate of generation of available potential energy",
                   "m^2 s^-3", ("l", "k"), {}),
}
QG_ONLY = {
    "APEflux": ("spectral flux of available potential energy", "m^2 s^-3", ("l", "k"),
                {"sums_with": ["paramspec_APEflux"]}),
    "KEflux": ("spectral flux of kinetic energy", "m^2 s^-3", ("l", "k"),
               {"sums_with": ["paramspec_KEflux"]}),
    "APEgen": ("total available potential energy generation", "m^2 s^-3", ("time",), {}),
}
LAYERED_ONLY = {
    "KEspec_modal": ("modal kinetic energy spectra", "m^2 s^-2", ("lev", "l", "k"), {}),
    "PEspec_modal": ("modal potential energy spectra", "m^2 s^-2", ("lev_mid", "l", "k"), {}),
    "APEspec": ("available potential energy spectrum", "m^2 s^-2", ("l", "k"), {}),
    "KEflux_div": ("spectral divergence of flux of kinetic energy", "m^2 s^-3", ("l", "k"),
                   {"sums_with": ["paramspec_KEflux"]}),
    "APEflux_div": ("spectral divergence of flux of available potential energy", "m^2 s^-3",
                    ("l", "k"), {"sums_with": ["paramspec_APEflux"]}),
}


def param_block(nz, Hi, H, S=None, del1=0.0, del2=0.0, rd=None, U=0.0, Vbg=None, Qx=None,
                f2gpi=None, pmodes=None, kdi2=None):
    """Per-member constants of the device diagnostics, in the layout of
    qgk_diag_set_params: Hi[nz], H, S[nz][nz], del1, del2, rd^-2, U1-U2,
    Vbg[nz], Qx[nz], f2gpi[nz], pmodes^-1 [nz][nz], kdi2[nz]."""
    def vec(v):
        return np.broadcast_to(np.asarray(0.0 if v is None else v,
                                          dtype=np.float64).ravel(), (nz,))
    S = np.zeros((nz, nz)) if S is None else np.asarray(S, dtype=np.float64).reshape(nz, nz)
    f2 = np.zeros(nz)
    if f2gpi is not None:
        f = np.asarray(f2gpi, dtype=np.float64).ravel()[:nz]
        f2[:f.size] = f
    pinv = np.linalg.inv(pmodes) if pmodes is not None else np.eye(nz)
    rd2inv = float(rd) ** -2 if rd else 0.0
    return np.concatenate([vec(Hi), [float(H)], S.ravel(),
                           [float(del1), float(del2), rd2inv, float(U)],
                           vec(Vbg), vec(Qx), f2, pinv.ravel(), vec(kdi2)])


class DeviceDiagnostic(dict):
    """A registry entry whose running sum ('value') is held on the device.

    It reads like pyqg's entry dictionaries: 'value' in entry is true once
    something was accumulated, entry['value'] is the sum (fetched from the
    device), entry['count'] the number of accumulations."""

    def __init__(self, model, name, **kw):
        super().__init__(**kw)
        self._model = model
        self._name = name

    def __contains__(self, key):
        if key == "value":
            return dict.__getitem__(self, "count") > 0
        return dict.__contains__(self, key)

    def __getitem__(self, key):
        if key == "value":
            if dict.__getitem__(self, "count") == 0:
                raise KeyError("value")
            return self._model._read_device_diagnostic(self._name)
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        return self[key] if key in self else default


class DiagnosticsMixin:
    """Registry and accumulation (pyqg/model.py:577-580, 630-822)."""

    _device_tables = (CORE,)

    def _initialize_diagnostics(self, diagnostics_list):
        self.diagnostics = dict()
        self._initialize_core_diagnostics()
        self._initialize_model_diagnostics()
        if diagnostics_list == 'all':
            pass
        elif diagnostics_list == 'none':
            self.set_active_diagnostics([])
        else:
            self.set_active_diagnostics(diagnostics_list)

    def _initialize_core_diagnostics(self):
        self._add_device_diagnostics(CORE)

    def _initialize_model_diagnostics(self):
        pass

    def _add_device_diagnostics(self, table):
        for name, (desc, units, dims, extra) in table.items():
            entry = DeviceDiagnostic(self, name, description=desc, units=units, dims=dims,
                                     active=True, count=0,
                                     function=(lambda m, n=name: m._device_diagnostic_now(n)))
            entry.update(extra)
            self.diagnostics[name] = entry

    def set_active_diagnostics(self, diagnostics_list):
        """Only the named diagnostics are accumulated."""
        for d in self.diagnostics:
            self.diagnostics[d]['active'] = d in diagnostics_list

    def add_diagnostic(self, diag_name, description=None, function=None, units=None, dims=None,
                       **kw):
        """A host-evaluated diagnostic (model.py:770-789)."""
        assert hasattr(function, '__call__')
        assert isinstance(diag_name, str)
        self.diagnostics[diag_name] = {'description': description, 'units': units,
                                       'dims': dims, 'active': True, 'count': 0,
                                       'function': function}
        self.diagnostics[diag_name].update(**kw)

    def describe_diagnostics(self):
        """Print a human-readable summary of the available diagnostics."""
        diag_names = sorted(self.diagnostics.keys())
        print('NAME               | DESCRIPTION')
        print(80 * '-')
        for k in diag_names:
            print('{:<10} | {:<54}'.format(k, self.diagnostics[k]['description']))

    def _diagnostics_due(self, t, tc):
        return (t >= self.dt) and (t >= self.tavestart) and (tc % self.taveints == 0)

    def _calc_diagnostics(self):
        if self._diagnostics_due(self.t, self.tc):
            self._increment_diagnostics()

    def _increment_diagnostics(self):
        device = [n for n, d in self.diagnostics.items()
                  if d['active'] and isinstance(d, DeviceDiagnostic)]
        host = [n for n, d in self.diagnostics.items()
                if d['active'] and not isinstance(d, DeviceDiagnostic)]
        if device:
            self._device_accumulate(device, instant=False)
            for n in device:
                self.diagnostics[n]['count'] += 1
        if host:
            self._calc_derived_fields()
            for n in host:
                d = self.diagnostics[n]
                res = d['function'](self)
                if d['count'] == 0:
                    d['value'] = res
                else:
                    d['value'] += res
                d['count'] += 1

    def get_diagnostic(self, dname):
        if 'value' not in self.diagnostics[dname]:
            raise DiagnosticNotFilledError(dname)
        return self.diagnostics[dname]['value'] / self.diagnostics[dname]['count']

    def _calc_derived_fields(self):
        pass

    # ------------------------------------------------------------- device side
    _DIAG_PARAM_INPUTS = ('Hi', 'H', 'S', 'del1', 'del2', 'rd', 'U', 'Vbg', 'Qx', 'f2gpi',
                          'pmodes', 'kdi2')

    def _diagnostic_params(self):
        """The per-member constant block of qgk_diag_set_params, rebuilt only
        when one of the model attributes it is made of changes."""
        key = tuple(np.asarray(v).tobytes() if v is not None else None
                    for v in (getattr(self, a, None) for a in self._DIAG_PARAM_INPUTS))
        cached = self.__dict__.get('_diag_param_cache')
        if cached is not None and cached[0] == key:
            return cached[1]
        nz = self.nz
        block = param_block(nz, self.Hi, self.H, S=getattr(self, 'S', None),
                            del1=getattr(self, 'del1', 0.0), del2=getattr(self, 'del2', 0.0),
                            rd=getattr(self, 'rd', None),
                            U=float(getattr(self, 'U', 0.0)) if nz == 2 else 0.0,
                            Vbg=getattr(self, 'Vbg', None), Qx=getattr(self, 'Qx', None),
                            f2gpi=getattr(self, 'f2gpi', None),
                            pmodes=getattr(self, 'pmodes', None),
                            kdi2=getattr(self, 'kdi2', None))
        self.__dict__['_diag_param_cache'] = (key, block)
        return block

    def _device_accumulate(self, names, instant):
        self._flush()
        b = self._bucket
        b.diag_set_params(0, self._diagnostic_params())
        weights, _ = C.ab_weights(self.ablevel, self.dt)
        b.diag_accumulate(names, self.dt, weights, instant=instant)

    def _read_device_diagnostic(self, name, instant=False):
        v = self._bucket.diag_read(name, member=0, instant=instant)
        if name in ("EKEdiss", "APEgen"):
            return float(v[0]) if np.ndim(v) else float(v)
        return v

    def _device_diagnostic_now(self, name):
        """Instantaneous value of a built-in diagnostic from the current fields."""
        self._device_accumulate([name], instant=True)
        return self._read_device_diagnostic(name, instant=True)
