"""raw-physics on MI355X: package placeholder (filled in by the product build)."""
