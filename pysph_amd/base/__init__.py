"""Particle storage, smoothing kernel and neighbour search (pysph.base)."""
