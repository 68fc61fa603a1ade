"""MI355X (gfx950) execution path for MiniLake's data-parallel core."""
