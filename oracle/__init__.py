"""The CPU oracle of the particle-grid hot path: test infrastructure only (see oracle.py)."""
