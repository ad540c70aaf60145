"""MI355X-native IMSRG(2) scalar-commutator / BCH / Magnus path
(drop-in for the commutator module of n-parzuchowski/EOM-IMSRG).

All arithmetic runs in libimsrg_hip.so (HIP, gfx950); importing the package
fails if the library has not been built.
"""
from . import _lib
from .basis import Basis
from .operators import CCMat, CCWork, SqOp
from .orbits import hk_orbits, oxygen16_holes
from . import commutators, synth
from .magnus import Flow, REF_TOL
from .batch import Batch
from . import adjoint, interaction, tune, tensor
from .adjoint import flow_energy

__all__ = ["Basis", "SqOp", "CCMat", "CCWork", "Flow", "REF_TOL", "commutators", "synth", "hk_orbits",
           "oxygen16_holes", "adjoint", "flow_energy", "tune", "Batch", "interaction", "tensor", "_lib"]
