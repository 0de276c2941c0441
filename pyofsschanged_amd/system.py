"""System, as pyofss/system.py:44-243: a list of modules called in turn with
(domain, field). Fibre modules run their propagation in the HIP engine."""

import numpy as np
import pandas as pd

from .domain import Domain
from .modules.fibre import Fibre


def field_save(field, filename="field_out"):
    """system.py:44-45"""
    np.savez_compressed(filename, field=field)


def field_load(filename="field_out"):
    """system.py:48-58"""
    if filename.endswith(".npz"):
        return np.load(filename)["field"]
    if filename.endswith(".npy"):
        return np.load(filename)
    try:
        return np.load(filename + ".npz")["field"]
    except FileNotFoundError:
        return np.load(filename + ".npy")


class System:
    """system.py:61-243 (run loop and result frames)."""

    def __init__(self, domain=None, field=None, charact_dir=None):
        self.domain = domain if domain is not None else Domain()
        self.field = None
        self.fields = None
        self.modules = None
        self.clear(remove_modules=True)
        if field is not None:
            self.field = field
        self.df_results = None
        self.z_curr = 0
        self.charact_dir = charact_dir

    def clear(self, remove_modules=False):
        """system.py:117-130"""
        if self.domain.channels > 1:
            self.field = [np.zeros([self.domain.total_samples], complex) for _ in range(self.domain.channels)]
        else:
            self.field = np.zeros([self.domain.total_samples], complex)
        self.fields = {}
        if remove_modules:
            self.modules = []

    def add(self, module):
        self.modules.append(module)

    def add_modules(self, module_list):
        for module in module_list:
            self.modules.append(module)

    def __getitem__(self, module_name):
        for module in self.modules:
            if module.name == module_name:
                return module

    def __setitem__(self, module_name, new_module):
        for index, module in enumerate(self.modules):
            if module.name == module_name:
                self.modules[index] = new_module
                return
        raise Exception("Tried to modify non-existing module in system")

    def update_result_df(self, obj):
        """system.py:187-202"""
        if type(obj) is Fibre and obj.stepper.storage.z:
            df_new = obj.stepper.storage.get_df_result(self.z_curr)
            self.z_curr = df_new.index.get_level_values("z [mm]").values[-1]
            self.df_results = df_new if self.df_results is None else pd.concat([self.df_results, df_new])

    def run(self):
        """system.py:220-243"""
        for module in self.modules:
            print(module.name)
            self.field = module(self.domain, self.field)
            self.fields[module.name] = self.field
            self.update_result_df(module)
