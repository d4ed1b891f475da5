"""MI355X (gfx950) engine for WarpGBM's training and inference hot path."""
from .core import WarpGBM

__all__ = ["WarpGBM"]
