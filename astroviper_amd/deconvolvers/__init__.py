"""Deconvolvers on the GPU (mirror of core/imaging/deconvolvers)."""
