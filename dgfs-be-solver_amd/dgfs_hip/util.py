"""Small helpers with the reference's semantics (frfs/util.py:140-163)."""
import itertools as it


def subclasses(cls, just_leaf=False):
    sc = cls.__subclasses__()
    ssc = [g for s in sc for g in subclasses(s, just_leaf)]
    return [s for s in sc if not just_leaf or not s.__subclasses__()] + ssc


def subclass_where(cls, **kwargs):
    """First subclass whose class attribute matches (frfs/util.py:151-159)."""
    k, v = next(iter(kwargs.items()))
    for s in subclasses(cls):
        if hasattr(s, k) and getattr(s, k) == v:
            return s
    raise KeyError(f"No subclasses of {cls.__name__} with cls.{k} == '{v}'")


def ndrange(*args):
    return it.product(*map(range, args))
