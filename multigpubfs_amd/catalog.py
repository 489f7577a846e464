"""Named CayleyPy constructors (multigpubfs/catalog.py:3-13): "Family.method"
with JSON-able args/kwargs -> GraphDefinition through from_cayleypy.  CayleyPy
is imported only when a catalog graph is requested."""
from __future__ import annotations

from .graph_definition import GraphDefinition, from_cayleypy

FAMILIES = ("PermutationGroups", "MatrixGroups", "Puzzles")


def from_catalog(name, *, args=None, kwargs=None) -> GraphDefinition:
    args = [] if args is None else args
    kwargs = {} if kwargs is None else kwargs
    if not isinstance(name, str) or name.count(".") != 1:
        raise ValueError("INVALID_CATALOG_CONSTRUCTOR")
    family, method = name.split(".")
    if family not in FAMILIES or not method.isidentifier() or method.startswith("_"):
        raise ValueError("INVALID_CATALOG_CONSTRUCTOR")
    if (not isinstance(args, (list, tuple)) or not isinstance(kwargs, dict)
            or any(not isinstance(k, str) or not k.isidentifier() or k.startswith("_") for k in kwargs)):
        raise ValueError("INVALID_CATALOG_ARGUMENTS")
    try:
        import cayleypy
    except ImportError as error:
        raise RuntimeError("CAYLEYPY_OPTIONAL_DEPENDENCY_REQUIRED: install cayleypy") from error
    constructor = getattr(getattr(cayleypy, family, None), method, None)
    if not callable(constructor):
        raise ValueError("UNKNOWN_CATALOG_CONSTRUCTOR")
    return from_cayleypy(constructor(*args, **kwargs))
