"""qaintensor: MI355X host package mirroring Qaintensor.jl's tensor-network
API over libqtn (hand-written gfx950 kernels behind a C ABI)."""
from ._lib import (Context, QtnError, QtnLibraryError, default_context,
                   load as load_library, lib_path)
from . import circuits
from .device import DeviceMPS, FORM_PLAIN, FORM_CHOSEN

__version__ = "0.1.0"
