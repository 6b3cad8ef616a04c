"""msplat_amd: MI355X-native 3D Gaussian Splatting training engine.

The public API mirrors the reference's `msplat` Python package (Dataset, GaussianTrainer,
TrainingConfig, TrainingStats, sync, cleanup) over libmsplat_core.so's C API. Importing
the package does not load the library; the first call does, and fails loudly if the
in-tree build is missing.
"""
from .core import Dataset, GaussianTrainer, TrainingConfig, TrainingStats, Viewer, cleanup, sync

__all__ = ["Dataset", "GaussianTrainer", "TrainingConfig", "TrainingStats", "Viewer", "sync", "cleanup"]
