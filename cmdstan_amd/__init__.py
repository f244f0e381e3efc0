"""cmdstan_amd: CmdStan's NUTS sampling path as batched HIP kernels for MI355X.

The product is the C-ABI library ``libcmdstan_amd.so`` (include/cmdstan_amd.h);
this package is its Python front end.
"""
from ._lib import LIB_PATH, lib, last_error  # noqa: F401
from .sampler import SAMPLER_COLUMNS, Data, Fit, Flow, Model, SamplerError, data_from_json  # noqa: F401

__all__ = ["Model", "Data", "Fit", "SamplerError", "SAMPLER_COLUMNS", "data_from_json", "lib", "LIB_PATH"]
