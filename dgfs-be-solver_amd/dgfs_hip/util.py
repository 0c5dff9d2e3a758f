"""Plug-in lookup by a class attribute.

The solver picks its scattering model, initial condition and boundary
classes by a string stored on the class (``scattering_model``, ``model``,
``type``), the selection contract of frfs/util.py:151-159 as used by
frfs/solvers/dgfs/system.py:69-72.  The lookup here walks the whole
inheritance tree below the base, nearest classes first, and returns the
first class that defines the attribute with the requested value.
"""
from collections import deque


def find_plugin(base, attr, value):
    """Class below ``base`` whose attribute ``attr`` equals ``value``.

    Classes are visited level by level (direct subclasses first), each once;
    a ``KeyError`` naming the base, the attribute and the value is raised
    when nothing matches, as a misspelt configuration key must fail loudly."""
    seen = set()
    todo = deque(base.__subclasses__())
    while todo:
        cls = todo.popleft()
        if cls in seen:
            continue
        seen.add(cls)
        if getattr(cls, attr, None) == value:
            return cls
        todo.extend(cls.__subclasses__())
    known = sorted({str(getattr(c, attr)) for c in seen if getattr(c, attr, None) is not None})
    raise KeyError(f"{base.__name__}: no plug-in with {attr} = {value!r} (known: {', '.join(known)})")


def subclass_where(cls, **kwargs):
    """``subclass_where(Base, attr=value)``: the reference's call form."""
    if len(kwargs) != 1:
        raise TypeError("subclass_where takes exactly one attr=value pair")
    (attr, value), = kwargs.items()
    return find_plugin(cls, attr, value)
