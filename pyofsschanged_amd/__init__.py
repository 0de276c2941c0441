"""MI355X (gfx950) engine for pyofss's fibre stepping path.

The public names follow pyofss/__init__.py:25-92 for the stepping path, so a
pyofss script built from Domain, System, Sech/Gaussian and Fibre runs
unchanged; each Fibre propagation runs in hand-written HIP kernels through
the C ABI of include/gnlse.h.
"""

import numpy as np  # noqa: F401  (pyofss re-exports numpy the same way)

from .domain import (Domain, align_with_nu_grid, dlambda_to_dnu, dnu_to_dlambda, lambda_to_nu, lambda_to_omega,
                     nu_to_lambda, nu_to_omega, omega_to_lambda, omega_to_nu)
from .field import (chirp, energy, fft, fftshift, get_downsampled, get_duration, get_duration_spec, ifft, ifftshift,
                    loss_infrared_db, loss_rayleigh_db, phase, spectral_power, spectrum_width_params,
                    temporal_power)
from .modules.amplifier import Amplifier, Amplifier2LevelModel
from .modules.fibre import Fibre
from .modules.gaussian import Gaussian
from .modules.linearity import (Linearity, convert_alpha_to_dB, convert_alpha_to_linear,
                                convert_dispersion_to_engineering, convert_dispersion_to_physical)
from .modules.nonlinearity import Nonlinearity, calculate_gamma, calculate_raman_term, calculate_raman_term_silica
from .modules.opencl_fibre import OpenclFibre, OpenclProgramm
from .modules.sech import Sech
from .modules.solver import Solver
from .modules.stepper import (MaximumStepsAllocatedError, SmallStepSizeError, Stepper, SuitableStepSizeError)
from .modules.storage import Storage
from .system import System, field_load, field_save

__all__ = [n for n in dir() if not n.startswith("_")]
