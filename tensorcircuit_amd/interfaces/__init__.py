"""Interfaces wrapping quantum functions for torch callers
(the reference's interfaces/ package, torch part)."""
from .torch import pytorch_interface, torch_interface  # noqa: F401
