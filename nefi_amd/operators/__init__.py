from .base import Fields, Operator

__all__ = ["Fields", "Operator"]
