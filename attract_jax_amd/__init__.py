"""MI355X-native ATTRACT rigid-body docking scoring and minimization.

Python mirror of the reference entry points (sjdv1982/attract-jax) over the C-ABI
of libattract_amd.so (include/attract_amd.h), whose HIP kernels do all compute.
"""
from ._lib import (ATJ_GRAD_EXACT, ATJ_LAYOUT_CORNER, ATJ_LAYOUT_PLAIN, ATJ_NB_PLATEAU, ATJ_POSE_QT, ATJ_TERM_GRID,
                   ATJ_TERM_NB, AttractError)
from .grid import GridFormatError, pad_grid, read_grid, read_grid2, write_grid
from .device import DeviceGrid, ForceField, Ligand, Receptor
from .score import (ScoringSystem, main, main_quat, neighbour_energy, potential_energrad, potential_energy,
                    value_and_grad_main, value_and_grad_quat)
from .minimize import BatchPlan, PoseState, minimize, minimize_plain, PlainGraph

__all__ = [
    "read_grid", "read_grid2", "pad_grid", "write_grid", "GridFormatError", "DeviceGrid", "ForceField", "Ligand",
    "Receptor", "ScoringSystem", "potential_energrad", "potential_energy", "neighbour_energy", "main",
    "value_and_grad_main", "main_quat", "value_and_grad_quat", "BatchPlan", "PoseState", "minimize",
    "minimize_plain", "PlainGraph", "AttractError", "ATJ_TERM_GRID", "ATJ_TERM_NB", "ATJ_NB_PLATEAU", "ATJ_GRAD_EXACT",
    "ATJ_POSE_QT", "ATJ_LAYOUT_PLAIN", "ATJ_LAYOUT_CORNER",
]
