"""pycd_amd: PyCD's kinetic Monte Carlo hot path on AMD Instinct MI355X.

Public API mirrors PyCD.core (System, Run, Analysis and their inputs).
"""
from .core import (Material, Neighbors, System, Run, Analysis, ReturnValues,
                   least_squares_slope)

__all__ = ["Material", "Neighbors", "System", "Run", "Analysis",
           "ReturnValues", "least_squares_slope"]
