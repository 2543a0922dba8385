"""Domain record restated from jubik/data.py:19-29 (TEST INFRASTRUCTURE ONLY)."""
from typing import NamedTuple


class Domain(NamedTuple):
    """Grid shape and pixel distances, as jubik.Domain."""

    shape: tuple
    distances: tuple
