"""muscade_amd — MI355X element-evaluation-and-assembly path of Muscade.jl.

Host-side mirror of Muscade's model/element/solver API for this path, calling
hand-written HIP kernels through the C ABI of include/muscade_hip.h.
"""
from ._lib import LIBPATH, lib
from .model import (DofGroup, Disassembler, Model, MuscadeError, State, addelement, addnode, allAdofs, allUdofs,
                    allXdofs, allXUdofs, allΛdofs, allΛXUAdofs, allΛXUdofs, decrement, getdof, getdofs,
                    gradientpartition, gradientstructure, indexedstate, initialize, nodofs, selecteddofs, setdof,
                    setscale)
from .elements import (Acost, AxisymmetricBarCrossSection, Bar3D, Bar3DU, BeamCrossSection, DofCost, DofLoad,
                       EulerBeam3D, EulerBeam3DU, EulerBeam3DA, EulerBeam3DAU, Hold, SingleAcost, SingleDofCost, SingleUdof,
                       StrainGaugeOnEulerBeam3D, ElementCost, ElementConstraint, DofConstraint, DofMotion, ResidualElement, equal,
                       positive, off)
from .assemble import AssemblyDirect, AssemblySweepX, DeviceContext, Newmarkβcoefficients
from .sweepx import SweepX, solve
from .directxua import AllStepsGradient, DirectXUA
from .hostcost import Dual2, exp10
from .results import getresult, request
from .adjoint import LimitStateGradient

__all__ = ["LIBPATH", "lib", "Model", "MuscadeError", "State", "Disassembler", "addnode", "addelement", "initialize",
           "setscale", "getdof", "setdof", "DofGroup", "allXdofs", "allΛdofs", "allUdofs", "allAdofs",
           "allXUdofs", "allΛXUdofs", "allΛXUAdofs", "nodofs", "selecteddofs", "indexedstate", "gradientstructure",
           "gradientpartition", "decrement", "getdofs", "BeamCrossSection", "EulerBeam3D", "EulerBeam3DU", "EulerBeam3DA", "EulerBeam3DAU", "Hold", "DofLoad",
           "AxisymmetricBarCrossSection", "Bar3D", "Bar3DU",
           "DofCost", "SingleDofCost", "SingleUdof", "Acost", "SingleAcost", "DirectXUA", "Dual2", "exp10",
           "StrainGaugeOnEulerBeam3D", "ResidualElement", "ElementCost", "ElementConstraint", "DofConstraint", "DofMotion", "equal", "positive", "off",
           "getresult", "request", "LimitStateGradient", "AllStepsGradient",
           "AssemblyDirect", "AssemblySweepX", "DeviceContext", "Newmarkβcoefficients", "SweepX", "solve"]
