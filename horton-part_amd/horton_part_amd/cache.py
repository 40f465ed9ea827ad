"""Result store and run-once decorator with the reference's semantics.

Restates ``core/cache.py`` (``JustOnceClass`` 34-63, ``just_once`` 65-80,
``Cache`` 160-399): ``load(key, alloc=...)`` returns ``(array, new)`` and
reuses arrays in place, ``dump`` stores values with tags, ``clear`` empties
arrays (keeping them allocated) unless ``dealloc`` is given, ``iterkeys``
filters on tags.  Added here: lazy entries (``add_lazy``) whose value is
computed on first access, so that per-point device results such as the AIM
weights are only copied to the host when a caller asks for them.
"""

import time

import numpy as np

__all__ = ["JustOnceClass", "just_once", "Cache"]


class JustOnceClass:
    """Base class for objects whose ``just_once`` methods run one time."""

    def __init__(self):
        self._done_just_once = set()
        self.time_usage = {}

    def __clear__(self):
        self.clear()

    def clear(self):
        self._done_just_once = set()


def just_once(fn):
    """Run ``fn(instance)`` only on its first call; record its wall time."""

    def wrapper(instance):
        if not hasattr(instance, "_done_just_once"):
            raise TypeError("Missing hidden _done_just_once. Forgot to call JustOnceClass.__init__()?")
        if fn.__name__ in instance._done_just_once:
            return
        t0 = time.time()
        fn(instance)
        instance._done_just_once.add(fn.__name__)
        instance.time_usage[fn.__name__] = time.time() - t0

    wrapper.__doc__ = fn.__doc__
    wrapper.__name__ = fn.__name__
    return wrapper


def _shape(alloc):
    if not hasattr(alloc, "__len__"):
        alloc = (alloc,)
    if len(alloc) == 0:
        raise TypeError("Alloc can not be an empty list")
    return tuple(alloc)


def _key(key):
    if hasattr(key, "__len__") and len(key) == 0:
        raise TypeError("At least one argument needed to specify a key.")
    while isinstance(key, tuple) and len(key) == 1:
        key = key[0]
    return key


class _Item:
    __slots__ = ("value", "valid", "tags")

    def __init__(self, value, tags):
        self.value = value
        self.valid = True
        self.tags = set(tags) if tags else set()


class Cache:
    """Named store of computed results (dictionary-like)."""

    def __init__(self):
        self._store = {}
        self._lazy = {}

    def add_lazy(self, key, producer, tags=None):
        """Register ``producer()`` to compute ``key`` on first access."""
        self._lazy[_key(key)] = (producer, tags)

    def _materialize(self, key):
        if key in self._lazy and (key not in self._store or not self._store[key].valid):
            producer, tags = self._lazy[key]
            self._store[key] = _Item(producer(), tags)

    def load(self, *key, alloc=None, default=None, tags=None, **kwargs):
        if kwargs:
            raise TypeError(f"Unknown optional arguments: {list(kwargs)}")
        has_default = default is not None
        key = _key(key)
        if alloc is not None and has_default:
            raise TypeError("The optional arguments alloc and default can not be used at the same time.")
        if tags is not None and alloc is None:
            raise TypeError("The tags argument is only allowed when the alloc argument is present.")
        self._materialize(key)
        item = self._store.get(key)
        if alloc is not None:
            shape = _shape(alloc)
            if item is None or not item.valid:
                if item is not None and isinstance(item.value, np.ndarray) and item.value.shape == shape:
                    item.valid = True
                    item.tags = set(tags) if tags else set()
                else:
                    item = _Item(np.zeros(shape), tags)
                    self._store[key] = item
                return item.value, True
            if not (isinstance(item.value, np.ndarray) and item.value.shape == shape):
                raise TypeError("The stored item does not match the given alloc.")
            if (set(tags) if tags else set()) != item.tags:
                raise ValueError("Tags do not match.")
            return item.value, False
        if item is None or not item.valid:
            if has_default:
                return default
            raise KeyError(key)
        return item.value

    def dump(self, *args, tags=None):
        if len(args) < 2:
            raise TypeError("At least two arguments are required: key1 and value.")
        key = _key(args[:-1])
        value = args[-1]
        item = self._store.get(key)
        if item is not None and item.valid and isinstance(item.value, np.ndarray) and isinstance(value, np.ndarray):
            if item.value.shape == value.shape and item.value.dtype == value.dtype:
                item.value[:] = value
                item.tags = set(tags) if tags else set()
                return
        self._store[key] = _Item(value, tags)

    def __contains__(self, key):
        key = _key(key)
        if key in self._lazy:
            return True
        item = self._store.get(key)
        return item is not None and item.valid

    def __len__(self):
        return sum(1 for k in self.iterkeys())

    def iterkeys(self, tags=None):
        tags = set(tags) if tags else set()
        keys = set(self._lazy) | {k for k, v in self._store.items() if v.valid}
        for key in sorted(keys, key=str):
            if key in self._store and self._store[key].valid:
                item_tags = self._store[key].tags
            else:
                item_tags = set(self._lazy[key][1] or ())
            if not tags or item_tags & tags:
                yield key

    def iteritems(self, tags=None):
        for key in self.iterkeys(tags):
            yield key, self.load(key)

    def clear_item(self, *key, dealloc=False):
        key = _key(key)
        self._lazy.pop(key, None)
        item = self._store.get(key)
        if item is None:
            return
        if not dealloc and isinstance(item.value, np.ndarray):
            item.valid = False
            item.value[:] = 0.0
        else:
            del self._store[key]

    def clear(self, dealloc=False, tags=None):
        tags = set(tags) if tags else set()
        for key in list(set(self._store) | set(self._lazy)):
            item = self._store.get(key)
            item_tags = item.tags if item is not None else set(self._lazy[key][1] or ())
            if not tags or item_tags & tags:
                self.clear_item(key, dealloc=dealloc)
