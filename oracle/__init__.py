"""CPU oracle of the Holovibes per-batch compute path (test infrastructure only)."""
