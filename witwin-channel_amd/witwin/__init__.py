"""Namespace package holding ``witwin.core`` (scene model) and ``witwin.channel`` (solvers)."""
