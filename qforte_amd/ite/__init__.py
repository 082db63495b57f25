"""Quantum imaginary-time evolution."""
