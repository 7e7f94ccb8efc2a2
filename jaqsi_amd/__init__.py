"""jaqsi_amd: an MI355X-native simulation backend with the JAQSI API.

Drop-in surface (jaqsi/__init__.py:27-122): ``Gates``, ``Script``, the gate
library, noise channels and observables.  Simulation runs in hand-written
HIP kernels for gfx950 (libjaqsi_amd.so) behind a ctypes C-ABI.
"""

from jaqsi_amd.operations import (  # noqa: F401
    GateStructure,
    Hermitian,
    Operation,
    cdtype,
    enable_x64,
    x64_enabled,
)
from jaqsi_amd.tape import active_tape, copy_to_tape, recording, shift_and_append  # noqa: F401
from jaqsi_amd import gateset, noise  # noqa: F401
from jaqsi_amd.gateset import (  # noqa: F401
    CCX,
    CPhase,
    CRX,
    CRY,
    CRZ,
    CSWAP,
    CX,
    CY,
    CZ,
    RX,
    RXX,
    RY,
    RYY,
    RZ,
    RZX,
    RZZ,
    SWAP,
    Barrier,
    ControlledPauliRot,
    ControlledPhaseShift,
    DiagonalQubitUnitary,
    H,
    Id,
    PauliRot,
    PauliX,
    PauliY,
    PauliZ,
    Rot,
    S,
    build_parity_observable,
)
from jaqsi_amd.gates import Gates  # noqa: F401
from jaqsi_amd.unitary import UnitaryGates  # noqa: F401
from jaqsi_amd.script import Script  # noqa: F401

__version__ = "0.1.0"
