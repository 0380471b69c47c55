"""fc3: MI355X-native third-order force constants and modal coupling constants.

A drop-in for ForceConstants.jl's third-order pipeline (second_order, third_order,
mass_weight_third_order!, mcc3, save_second_order / save_third_order) backed by
hand-written HIP kernels for gfx950 in libfc3.so.  See api.py for the
reference-named entry points.
"""
from . import systems
from .api import (asr_satisfied, autotune_load, autotune_query, autotune_save, batch_graph_lanes,
                  batch_graph_nodes,
                  batch_third_order_mcc3, Context, DenseForceConstants, DeviceArray,
                  dynamical_matrix, format_float, get_modes, issymmetric,
                  k3_from_unique, mass_weight_third_order_, mcc3, mcc3_unique, unique_count,
                  unique_index,
                  mcc3_, mcc3_backward, mcc3_rrule, modes_device, save_second_order,
                  save_third_order, second_order, second_order_device, third_order,
                  third_order_device, third_order_mcc3, third_order_param_backward,
                  weighted_square_loss)
from .systems import (AnalyticalCalculator, AutoDiffCalculator, LJ, StillingerWeberSilicon,
                      SuperCellSystem, SWParams, diamond, fcc)

__all__ = [
    "asr_satisfied", "autotune_load", "autotune_query", "autotune_save", "batch_graph_lanes",
    "batch_graph_nodes",
    "batch_third_order_mcc3", "Context", "DenseForceConstants", "DeviceArray",
    "dynamical_matrix", "format_float", "get_modes", "issymmetric", "k3_from_unique", "mass_weight_third_order_", "mcc3",
    "mcc3_unique", "unique_count", "unique_index",
    "mcc3_", "mcc3_backward", "mcc3_rrule", "modes_device", "save_second_order",
    "save_third_order", "second_order", "second_order_device", "third_order",
    "third_order_device", "third_order_mcc3", "third_order_param_backward", "systems",
    "weighted_square_loss",
    "AnalyticalCalculator", "AutoDiffCalculator", "LJ", "StillingerWeberSilicon",
    "SuperCellSystem", "SWParams", "diamond", "fcc",
]
