"""SPH functions and the calc that evaluates them (pysph.sph)."""
