"""MI355X (gfx950) pseudospectral Navier-Stokes box: the rhs / rhsAdj / stepper
hot path of dnsbox as hand-written HIP kernels behind a C-ABI
(include/dnsbox.h), with a Python front-end over the same library."""
from ._lib import DnsboxError, load  # noqa: F401
from .box import Box, BoxGroup, default_params  # noqa: F401

__all__ = ["Box", "BoxGroup", "default_params", "DnsboxError", "load"]
