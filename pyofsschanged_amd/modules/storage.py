"""Trace storage, as pyofss/modules/storage.py:41-350.

Traces come back from the device as distinct arrays (the reference keeps a
reference to each stored field, storage.py:190-191). The per-trace
characteristics of update_characts (storage.py:194-209) of a whole
propagation are computed on the device in one launch (append_traces);
append keeps the reference's one-trace form for direct callers.
"""

import os
import warnings

import numpy as np
import pandas as pd

from .. import field
from ..field import (energy, get_duration, get_duration_spec, get_peaks, simpson_weights, spectral_power,
                     temporal_power)


class StorageError(Exception):
    pass


class DirExistenceError(StorageError):
    pass


class DifferentAxisError(StorageError):
    pass


class SavingWarning(Warning):
    pass


class InvalidArgumentError(StorageError):
    pass


def check_dir(dir_name):
    """storage.py:112-120"""
    if not os.path.exists(dir_name):
        os.makedirs(dir_name)
        print("Directory: ", dir_name, " is created!")
    elif not os.listdir(dir_name):
        print("Directory is empty")
    else:
        warnings.warn("Directory is not empty")


class Storage:
    """storage.py:130-350"""

    def __init__(self, dir=None, cycle=None, fibre_name=None, f=None, downsampling=None):
        self.dir = dir
        if dir is not None:
            check_dir(dir)
        self.f = f
        self.cycle = cycle
        self.fibre_name = fibre_name
        self.downsampling = downsampling
        self.plt_data = None
        self.domain = None
        self.As = []
        self.z = []
        self.step_sizes = []
        self.fft_total = 0
        self.energy_list = []
        self.max_power_list = []
        self.peaks_list = []
        self.duration_list = []
        self.spec_width_list = []
        self.l_nl_list = []
        self.l_d_list = []

    def __call__(self, domain):
        self.domain = domain

    def reset_array(self):
        self.As = []
        self.z = []

    @staticmethod
    def reset_fft_counter():
        """storage.py:174-177"""
        field.fft_counter = 0

    def store_current_fft_count(self):
        """storage.py:179-181"""
        self.fft_total = field.fft_counter

    def append(self, z, A):
        """storage.py:183-192"""
        self.z.append(z)
        self.As.append(A)
        self.update_characts(A)

    def append_traces(self, zs, As, ctx):
        """storage.py:183-209 for all traces of a propagation: z and field
        are stored as by append, and the characteristics (energy, maximum
        power, peaks, duration, spectral width, L_D, L_NL) of every trace come
        from one device launch on the context's grid. Differences from the
        reference: a trace without a spectral peak of prominence >= max/100
        gets a NaN spectral width (the reference fails on abs(None),
        field.py:278) and one without any temporal peak a NaN duration."""
        self.z.extend(zs)
        self.As.extend(As)
        self._ctx = ctx
        if self.domain is None or self.f is None or len(As) == 0:
            return
        field.fft_counter += len(As)  # spectral_power's field.fft per trace (storage.py:203)
        if np.ndim(As[0]) > 1:  # WDM traces: the reference's find_peaks fails on 2-channel powers
            return
        ch = ctx.trace_characteristics(np.stack(As), simpson_weights(self.domain.t * 1e-3))
        for k in range(len(As)):
            T0 = float(ch["duration"][k]) * self.domain.dt
            P0 = float(ch["max_power"][k])
            self.energy_list.append(float(ch["energy"][k]))
            self.max_power_list.append(P0)
            self.peaks_list.append(ch["peaks"][k])
            self.duration_list.append(T0)
            self.spec_width_list.append(float(ch["spec_width"][k]) * self.domain.dnu)
            L_D, L_NL = self.f.lengths(T0, P0)
            self.l_d_list.append(L_D)
            self.l_nl_list.append(L_NL)

    def update_characts(self, A):
        """storage.py:194-209 (skipped without a domain or length functions)."""
        if self.domain is None or self.f is None:
            return
        self.energy_list.append(energy(A, self.domain.t))
        P_t = temporal_power(A)
        P_nu = spectral_power(A)
        self.max_power_list.append(np.amax(P_t))
        self.peaks_list.append(get_peaks(P_t))
        self.duration_list.append(get_duration(P_t, self.domain.dt))
        self.spec_width_list.append(get_duration_spec(P_nu, self.domain.dnu))
        self.l_d_list.append(self.f.l_d(A))
        self.l_nl_list.append(self.f.l_nl(A))

    def save_all_storage_to_dir_as_df(self, save_power=True, channel=None):
        """storage.py:211-248"""
        if self.dir is None:
            warnings.warn("Nothing will be saved - the base fibre directory is not stated!", SavingWarning)
            return
        if save_power:
            for kind in ("temp", "spec"):
                d = os.path.join(self.dir, kind)
                check_dir(d)
                self.get_df(kind).to_csv(os.path.join(d, f"{self.fibre_name}.csv"))
        else:
            d = os.path.join(self.dir, "complex")
            check_dir(d)
            self.get_df("complex").to_csv(os.path.join(d, f"{self.fibre_name}.csv"))
        with open(os.path.join(self.dir, "current_info.txt"), "w") as f:
            f.write(f"current cycle: {self.cycle}, current fibre: {self.fibre_name}")

    def _index(self, z_curr):
        arr_z = np.array(self.z) * 10 ** 6 + z_curr  # mm
        if self.cycle and self.fibre_name is not None:
            return pd.MultiIndex.from_product([[self.cycle], [self.fibre_name], arr_z],
                                              names=["cycle", "fibre", "z [mm]"])
        return pd.MultiIndex.from_product([arr_z], names=["z [mm]"])

    def get_df(self, type="complex", z_curr=0, channel=None):
        """storage.py:250-268"""
        if type == "temp":
            _, y, _ = self.get_plot_data(is_temporal=True)
        elif type == "spec":
            _, y, _ = self.get_plot_data(is_temporal=False)
        elif type == "complex":
            y = self.As
        else:
            raise InvalidArgumentError(f"{type} is not a valid argument, type param can be 'temp', 'spec' or 'complex'")
        return pd.DataFrame(y, index=self._index(z_curr))

    def get_df_result(self, z_curr=0):
        """storage.py:270-302"""
        cols = ["Peak Power [W]", "Energy [nJ]", "Temp width [ps]", "Spec width [THz]", "Dispersion length [km]",
                "Nonlinear length [km]", "Peaks [idx]"]
        df = pd.DataFrame(index=self._index(z_curr), columns=cols)
        if len(self.max_power_list) == len(self.z):
            df["Peak Power [W]"] = self.max_power_list
            df["Energy [nJ]"] = self.energy_list
            df["Temp width [ps]"] = self.duration_list
            df["Spec width [THz]"] = self.spec_width_list
            df["Peaks [idx]"] = self.peaks_list
            df["Dispersion length [km]"] = self.l_d_list
            df["Nonlinear length [km]"] = self.l_nl_list
        return df

    def get_plot_data(self, is_temporal=True, reduced_range=None, normalised=False, channel=None):
        """storage.py:304-350. With down-sampling (and neither range
        reduction nor normalisation) the powers are formed and down-sampled
        on the device, both domains in one launch, and kept for the other
        domain's call."""
        ctx = getattr(self, "_ctx", None)
        if (self.downsampling is not None and reduced_range is None and not normalised and ctx is not None
                and len(self.As) > 0 and ctx.n == len(self.As[0])):
            key = (id(self.As), len(self.As), int(self.downsampling))
            if getattr(self, "_down_key", None) != key:
                from .._lib import DOWNSAMPLE_PYOFSS
                self._down = ctx.trace_powers(np.stack(self.As), int(self.downsampling), DOWNSAMPLE_PYOFSS)
                self._down_key = key
            if not is_temporal:
                field.fft_counter += len(self.As)  # spectral_power's field.fft per trace
            x = self.domain.t if is_temporal else self.domain.nu
            y = list(self._down[0] if is_temporal else self._down[1])
            self.plt_data = (x, y, self.z)
            return self.plt_data
        if is_temporal:
            x = self.domain.t
            power = [temporal_power(A) for A in self.As]
        else:
            x = self.domain.nu
            power = [spectral_power(A) for A in self.As]
        if normalised:
            factor = max(power[0])
            power = [p / factor for p in power]
        if reduced_range is not None:
            lo = np.abs(x - reduced_range[0]).argmin()
            hi = np.abs(x - reduced_range[1]).argmin()
            x = x[lo:hi + 1]
            power = [p[lo:hi + 1] for p in power]
        if self.downsampling is not None:
            power = [field.get_downsampled(p, self.downsampling) for p in power]
        self.plt_data = (x, power, self.z)
        return self.plt_data
