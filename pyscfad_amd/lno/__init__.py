"""Local-natural-orbital (LNO) fragment (T) on MI355X."""
