"""System, as pyofss/system.py:44-243: a list of modules called in turn with
(domain, field). Fibre modules run their propagation in the HIP engine."""

import os

import numpy as np
import pandas as pd

from .domain import Domain
from .modules.fibre import Fibre
from .modules.opencl_fibre import OpenclFibre
from .modules.storage import check_dir


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
    """system.py:61-243: run loop, per-fibre result frames (df_results), field
    frames per cycle (df_fields) and the lazily built trace frames
    (df_temp, df_spec, df_complex)."""

    def __init__(self, domain=None, field=None, charact_dir=None):
        self.domain = domain if domain is not None else Domain()
        self.field = None
        self.fields = None
        self.modules = None
        self.clear(remove_modules=True)
        if field is not None:
            self.field = field
        self.df_results = None
        self.df_fields = None
        self.df_type_dict = {"temp": None, "spec": None, "complex": None}
        self.z_curr = 0
        self.charact_dir = charact_dir
        if self.charact_dir is not None:
            check_dir(self.charact_dir)

    def _df(self, kind):
        if self.df_type_dict[kind] is None:
            self.init_df(kind)
        return self.df_type_dict[kind]

    @property
    def df_temp(self):
        """system.py:91-97"""
        return self._df("temp")

    @property
    def df_spec(self):
        """system.py:99-105"""
        return self._df("spec")

    @property
    def df_complex(self):
        """system.py:107-113"""
        return self._df("complex")

    def clear(self, remove_modules=False):
        """system.py:115-130"""
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

    def init_df(self, df_type="complex"):
        """system.py:153-172: concatenate the fibres' trace frames, each
        continuing the z axis of the previous one."""
        if df_type not in self.df_type_dict:
            raise ValueError()
        df = pd.DataFrame()
        z_curr = 0
        for obj in self.modules:
            df_new = None
            if type(obj) is Fibre:
                df_new = obj.stepper.storage.get_df(type=df_type, z_curr=z_curr)
            elif type(obj) is OpenclFibre:
                df_new = obj.get_df(type=df_type, z_curr=z_curr)
            if df_new is not None:
                if not df_new.empty:
                    z_curr = df_new.iloc[-1].name[-1]
                    df = pd.concat([df, df_new])
        self.df_type_dict[df_type] = df

    def save_df_to_csv(self, dir, df_type="complex", name="laser"):
        """system.py:175-184"""
        check_dir(dir)
        if df_type not in self.df_type_dict:
            raise ValueError()
        if self.df_type_dict[df_type] is None:
            self.init_df(df_type=df_type)
        self.df_type_dict[df_type].to_csv(os.path.join(dir, name + "_" + df_type + ".csv"))

    def update_laser_info(self, obj):
        """system.py:186-199. A fibre that stored no trace (traces=1) adds no
        rows; the reference fails there indexing the empty z axis."""
        df_new = None
        if type(obj) is Fibre and obj.stepper.storage.z:
            df_new = obj.stepper.storage.get_df_result(self.z_curr)
        elif type(obj) is OpenclFibre and obj.z_list:
            df_new = obj.get_df_result(self.z_curr)
        if df_new is not None:
            self.z_curr = df_new.index.get_level_values("z [mm]").values[-1]
            self.df_results = df_new if self.df_results is None else pd.concat([self.df_results, df_new])

    def update_result_df(self, obj):
        """system.py:201-202"""
        self.update_laser_info(obj)

    def save_result_df_to_scv(self, dir):
        """system.py:204-207"""
        check_dir(dir)
        if self.df_results is not None:
            self.df_results.to_csv(os.path.join(dir, "laser_info.csv"))

    def update_fields_df(self, df_field_new):
        """system.py:209-213"""
        self.df_fields = df_field_new if self.df_fields is None else pd.concat([self.df_fields, df_field_new])

    def save_fields_df_to_scv(self, dir):
        """system.py:215-218"""
        check_dir(dir)
        if self.df_fields is not None:
            self.df_fields.to_csv(os.path.join(dir, "fields.csv"))

    def run(self):
        """system.py:220-243"""
        for module in self.modules:
            print(module.name)
            self.field = module(self.domain, self.field)
            self.fields[module.name] = self.field
            self.update_result_df(module)
            if getattr(module, "cycle", None) and module.name is not None:
                index = pd.MultiIndex.from_product([[module.cycle], [module.name]], names=["cycle", "fibre"])
                self.update_fields_df(pd.DataFrame([self.field], index=index))
                if self.charact_dir is not None:
                    self.save_fields_df_to_scv(self.charact_dir)
            if module.name == "splitter_2" and self.charact_dir:
                print("new characts are saved")
                self.save_result_df_to_scv(self.charact_dir)
