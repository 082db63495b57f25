"""Classical solvers used by the drivers."""
