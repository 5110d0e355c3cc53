"""Host-side helpers: array geometry, validation, timing, CLI ranges, fills."""
