"""MI355X (gfx950) Brouwer–Zimmermann minimum-distance engine.

``qubitserf_amd.distfind`` mirrors the reference's ``qubitserf.distfind`` entry points for
the ``backend="gpu"`` path; ``qubitserf_amd.codeaut`` the ``codeaut_bz_collect`` low-weight
enumeration; ``qubitserf_amd.engine`` the plugin-level calls (one weight level at a time).
"""
__version__ = "0.1.0"
