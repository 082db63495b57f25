"""Hamiltonian sources: external JSON integrals and synthetic integrals."""
