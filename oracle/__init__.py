"""CPU oracle of the ray tracer: test infrastructure only (see oracle.py)."""
