"""qaintensor: MI355X host package mirroring Qaintensor.jl's tensor-network
API over libqtn (hand-written gfx950 kernels behind a C ABI)."""
from ._lib import (Context, QtnError, QtnLibraryError, default_context,
                   load as load_library, lib_path)
from . import autotune, circuits, learned
from .device import (DeviceMPO, DeviceMPS, FORM_PLAIN, FORM_CHOSEN,
                     apply_1q_multi, apply_layer_multi, run_circuits)
from .network import (ErrorException, GeneralTensorNetwork, Summation,
                      Tensor, TensorNetwork, contract, contract_rep,
                      is_power_two, shift_pair, shift_summation)
from .mps import (MPS, ClosedMPS, OpenMPS, PeriodicMPS, check_mps,
                  contract_svd, contract_svd_mps, permute_mps, switch)
from .mpo import (MPO, CircuitGate, apply_MPO, decompose, extend_MPO,
                  tensor_circuit)

__version__ = "0.1.0"
