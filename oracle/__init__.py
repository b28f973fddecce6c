"""The checker of the sector matvec / Krylov path (test infrastructure; see pyoracle.py)."""
