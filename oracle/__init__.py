"""CPU checkers of the level-set update path (test infrastructure only)."""
