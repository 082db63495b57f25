"""Algorithm base classes (drivers' common machinery)."""
