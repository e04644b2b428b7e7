"""Exact code distances on the gfx950 engine (mirror of ``qubitserf.distfind``)."""
from .api import (Session, available_backends, batch_distance, classical_distance, css_distance, device_memory,
                  last_batch_stats, last_trace, set_sound_bounds, sound_bounds, stabilizer_distance, version)

__all__ = ["css_distance", "classical_distance", "stabilizer_distance", "available_backends",
           "version", "last_trace", "batch_distance", "last_batch_stats", "Session", "set_sound_bounds",
           "sound_bounds", "device_memory"]
