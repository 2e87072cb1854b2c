"""Operations, integrators and the solver time loop (pysph.solver)."""
