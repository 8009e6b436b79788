"""Acoustic wave propagation (2-D propagator, 2-D and 3-D kernels), mirroring the import surface of
``geobrain.physics.wave`` (``geobrain/physics/wave/__init__.py``) for the
acoustic path: configs, result, model, survey, wavelet, PML boundary,
propagator and the ``kernels`` module."""

from .config import BoundaryConfig, GridConfig
from .result import WaveResult
from .models import AbstractModel, AcousticModel
from .wavelets import RickerWavelet, WaveletGenerator
from .boundary import bc_pml, bc_pml_3d, create_boundary
from .survey import Receiver, Source, Survey
from .propagators import AbstractPropagator, AcousticPropagator
from .utils import numpy2tensor, tensor2numpy
from . import kernels
from .kernels import acoustic_forward_kernel, acoustic_forward_kernel_3d, pad_model, pad_model_3d

__all__ = [
    'GridConfig', 'BoundaryConfig', 'WaveResult', 'AbstractModel', 'AcousticModel',
    'WaveletGenerator', 'RickerWavelet', 'bc_pml', 'bc_pml_3d', 'create_boundary', 'Source', 'Receiver',
    'Survey', 'AbstractPropagator', 'AcousticPropagator', 'numpy2tensor', 'tensor2numpy',
    'kernels', 'acoustic_forward_kernel', 'acoustic_forward_kernel_3d', 'pad_model',
    'pad_model_3d',
]
