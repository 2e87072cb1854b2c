"""Time-step policies (solver/time_step_functions.py)."""


class TimeStep:
    """Fixed time step: returns solver.dt (time_step_functions.py:3-6)."""

    def compute_time_step(self, solver):
        return solver.dt
