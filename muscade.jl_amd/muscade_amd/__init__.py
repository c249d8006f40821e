"""muscade_amd — MI355X element-evaluation-and-assembly path of Muscade.jl.

Host-side mirror of Muscade's model/element/solver API for this path, calling
hand-written HIP kernels through the C ABI of include/muscade_hip.h.
"""
from ._lib import LIBPATH, lib
from .model import (Disassembler, Model, MuscadeError, State, addelement, addnode, getdof, initialize, setdof,
                    setscale)
from .elements import (Acost, AxisymmetricBarCrossSection, Bar3D, Bar3DU, BeamCrossSection, DofCost, DofLoad,
                       EulerBeam3D, EulerBeam3DU, Hold, SingleAcost, SingleDofCost, SingleUdof,
                       StrainGaugeOnEulerBeam3D, ElementCost, ElementConstraint, DofConstraint, equal, positive,
                       off)
from .assemble import AssemblyDirect, AssemblySweepX, DeviceContext, Newmarkβcoefficients
from .sweepx import SweepX, solve
from .directxua import DirectXUA
from .hostcost import Dual2, exp10
from .results import getresult, request

__all__ = ["LIBPATH", "lib", "Model", "MuscadeError", "State", "Disassembler", "addnode", "addelement", "initialize",
           "setscale", "getdof", "setdof", "BeamCrossSection", "EulerBeam3D", "EulerBeam3DU", "Hold", "DofLoad",
           "AxisymmetricBarCrossSection", "Bar3D", "Bar3DU",
           "DofCost", "SingleDofCost", "SingleUdof", "Acost", "SingleAcost", "DirectXUA", "Dual2", "exp10",
           "StrainGaugeOnEulerBeam3D", "ElementCost", "ElementConstraint", "DofConstraint", "equal", "positive", "off",
           "getresult", "request",
           "AssemblyDirect", "AssemblySweepX", "DeviceContext", "Newmarkβcoefficients", "SweepX", "solve"]
