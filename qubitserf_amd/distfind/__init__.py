"""Exact code distances on the gfx950 engine (mirror of ``qubitserf.distfind``)."""
from .api import (available_backends, classical_distance, css_distance, last_trace,
                  stabilizer_distance, version)

__all__ = ["css_distance", "classical_distance", "stabilizer_distance", "available_backends",
           "version", "last_trace"]
