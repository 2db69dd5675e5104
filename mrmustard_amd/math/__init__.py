"""``mrmustard_amd.math``: the backend-manager singleton, as ``mrmustard.math``
is one (mrmustard/math/__init__.py:34)."""

from .backend_manager import BackendManager

math = BackendManager()

__all__ = ["BackendManager", "math"]
