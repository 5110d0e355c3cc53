"""Command line scripts (sweep collections, CSV analysis)."""
