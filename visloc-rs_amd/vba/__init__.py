"""Bundle adjustment on AMD MI355X: Python host side of libvba.so.

Drop-in for the Levenberg–Marquardt Schur-complement path of visloc-rs
(`BundleAdjustment::optimize`, `BaAccelerator`); see include/vba.h.
"""
from . import abi
from .abi import library
from .api import (BatchOutput, BundleProblem, Device, PointEliminationInput, PreparedBatch, PreparedProblem,
                  PrepareError, ServedResult,
                  Server, SolveOutput, Synthetic, camera, camera_centres, check, default_gnc_options, default_options,
                  run_solver, sensor, synth_params, synthesize)
from .io import IoError, Model

__all__ = [
    "abi", "library", "BatchOutput", "BundleProblem", "Device", "PointEliminationInput", "PreparedBatch", "PreparedProblem", "PrepareError",
    "ServedResult", "Server", "SolveOutput", "Synthetic", "camera", "camera_centres", "check", "default_gnc_options",
    "default_options",
    "run_solver", "sensor", "synth_params", "synthesize", "IoError", "Model",
]
