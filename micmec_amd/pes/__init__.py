"""Potential-energy surface: the force-field plugin API and its parts."""

from .mmff import (ForcePart, MicMecForceField, ForcePartMechanical, ForcePartPressure,
                   parse_type_params)

__all__ = ["ForcePart", "MicMecForceField", "ForcePartMechanical", "ForcePartPressure",
           "parse_type_params"]
