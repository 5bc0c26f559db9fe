"""Utilities: physical constants and geometry shapes."""

from . import constants, shape  # noqa: F401
