"""Imaging weights on the GPU (mirror of core/imaging/imaging_weighting)."""
