"""Blackframe XPU/HIP: MI355X wavefront path-tracing backend (host-side Python mirror)."""
from . import abi, scenes  # noqa: F401
