"""Integrators, thermostats, barostats and optimisers (micmec/sampling)."""

from .iterative import (Iterative, StateItem, AttributeStateItem, PosStateItem,
                        TemperatureStateItem, VolumeStateItem, DomainStateItem,
                        ConsErrStateItem, EPotContribStateItem, Hook)
from .verlet import VerletIntegrator, VerletHook, VerletScreenLog, ConsErrTracker
from .nvt import NHChain, NHCThermostat, LangevinThermostat, NHCAttributeStateItem
from .npt import TBCombination, MTKBarostat, LangevinBarostat, MTKAttributeStateItem
from .dof import DOF, CartesianDOF, BaseCellDOF, FullCellDOF, StrainCellDOF
from .opt import OptScreenLog, BaseOptimizer, CGOptimizer, ConjugateGradientSearch
from .trajectory import BaseHDF5Writer, HDF5Writer, XYZWriter

__all__ = ["Iterative", "StateItem", "AttributeStateItem", "PosStateItem",
           "TemperatureStateItem", "VolumeStateItem", "DomainStateItem", "ConsErrStateItem",
           "EPotContribStateItem", "Hook", "VerletIntegrator", "VerletHook", "VerletScreenLog",
           "ConsErrTracker", "NHChain", "NHCThermostat", "LangevinThermostat",
           "NHCAttributeStateItem", "TBCombination", "MTKBarostat", "LangevinBarostat",
           "MTKAttributeStateItem", "DOF", "CartesianDOF", "BaseCellDOF", "FullCellDOF",
           "StrainCellDOF",
           "OptScreenLog", "BaseOptimizer", "CGOptimizer", "ConjugateGradientSearch",
           "BaseHDF5Writer", "HDF5Writer", "XYZWriter"]
