"""Solver orchestration: padding + absorbing layer, engine inputs, launcher."""
