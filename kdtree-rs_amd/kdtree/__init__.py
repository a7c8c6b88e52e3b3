"""MI355X KD-tree: exact k-NN behind the reference's `kdtree.KDTree` API.

`from kdtree import KDTree` as with the reference (python/kdtree/__init__.py).
"""

from ._core import KDTree

__all__ = ["KDTree"]
