"""Domain record and the read-out container.

``Domain`` mirrors jubik/data.py:19-29.  ``Vector`` is the minimal
stand-in for ``nifty.re.Vector`` that the read-out returns
(jubik/response.py:110): a mapping of keys to 1-D arrays exposed as ``.tree``.
"""
from typing import NamedTuple


class Domain(NamedTuple):
    """Grid shape and pixel distances (jubik.Domain)."""

    shape: tuple
    distances: tuple


def is_domain(obj):
    """True for this Domain and for any NamedTuple called Domain with the
    same fields (so jubik.Domain instances pass unchanged)."""
    if isinstance(obj, Domain):
        return True
    return (type(obj).__name__ == "Domain" and isinstance(obj, tuple)
            and getattr(obj, "_fields", None) == ("shape", "distances"))


class Vector:
    """Keyed collection of read-out arrays, as ``nifty.re.Vector``."""

    def __init__(self, tree):
        self.tree = dict(tree)

    def __getitem__(self, key):
        return self.tree[key]

    def __iter__(self):
        return iter(self.tree)

    def __len__(self):
        return len(self.tree)

    def keys(self):
        return self.tree.keys()

    def items(self):
        return self.tree.items()

    def __repr__(self):
        return f"Vector({list(self.tree)})"
