"""vtp - MI355X particle-fluid step (PBF, DFSPH, IMM, IMM-CT) behind a C ABI.

Python is a thin layer over libvtp.so (include/vtp.h): ctypes bindings,
scene generators, and the torch.autograd.Function of the PBF adjoint.
"""
from ._abi import (DFSPH, IMM, IMMCT, MAT_BOUNDARY, MAT_FLUID, PBF, Context, VtpError, VtpParams,
                   VtpStateHeader, config_template, load_library, object_config_template, params_default,
                   params_from_json, parse_state)
from . import scenes

__all__ = ["Context", "VtpError", "VtpParams", "VtpStateHeader", "parse_state", "config_template", "object_config_template", "load_library", "params_default", "params_from_json", "scenes",
           "PBF", "DFSPH", "IMM", "IMMCT", "MAT_FLUID", "MAT_BOUNDARY"]
