"""sfft_amd: MI355X-native sparse FFT behind the reference sFFT C API.

The product is ``libsfft_amd.so`` (C ABI in ``include/sfft.h``: host C++
plus hand-written gfx950 HIP kernels).  This package is its Python host
interface; see ``sfft_amd.sfft.sfft``.
"""
from .sfft import (FFTW_ESTIMATE, FFTW_MEASURE, V1_V2_INPUT_PARAMETERS, SfftError,
                   derive_params, design_filter, sfft)
from . import _lib

__all__ = ["sfft", "SfftError", "derive_params", "design_filter", "FFTW_ESTIMATE",
           "FFTW_MEASURE", "V1_V2_INPUT_PARAMETERS"]


def library_path():
    return _lib.LIB_PATH
