"""Base class of iterative algorithms, state items and hooks (micmec/sampling/iterative.py:42-222)."""

import copy

import numpy as np

__all__ = ["Iterative", "StateItem", "AttributeStateItem", "PosStateItem",
           "EPotContribStateItem", "ConsErrStateItem", "TemperatureStateItem",
           "VolumeStateItem", "DomainStateItem", "Hook"]


class Iterative(object):
    default_state = []
    log_name = "ITER"

    def __init__(self, mmf, state=None, hooks=None, counter0=0):
        self.mmf = mmf
        self.state_list = [item.copy() for item in self.default_state]
        if state is not None:
            self.state_list += state
        self.state = dict((item.key, item) for item in self.state_list)
        if hooks is None:
            self.hooks = []
        elif hasattr(hooks, "__len__"):
            self.hooks = hooks
        else:
            self.hooks = [hooks]
        self._add_default_hooks()
        self.counter0 = counter0
        self.counter = counter0
        self.initialize()

    def _add_default_hooks(self):
        pass

    def initialize(self):
        self.call_hooks()

    def _hooks_due(self, counter):
        return any(hook.expects_call(counter) for hook in self.hooks
                   if not getattr(hook, "_device_hook", False))

    def call_hooks(self):
        """Update all state items once, then call every hook that is due (iterative.py:77-86)."""
        state_updated = False
        for hook in self.hooks:
            if getattr(hook, "_device_hook", False):
                continue
            if hook.expects_call(self.counter):
                if not state_updated:
                    self._sync_for_hooks()
                    for item in self.state_list:
                        item.update(self)
                    state_updated = True
                hook(self)

    def _sync_for_hooks(self):
        """Bring host-visible attributes up to date before hooks read them."""

    def run(self, nsteps=None):
        if nsteps is None:
            while True:
                if self.propagate():
                    break
        else:
            for _ in range(nsteps):
                if self.propagate():
                    break
        self.finalize()

    def propagate(self):
        self.counter += 1
        self.call_hooks()

    def finalize(self):
        pass


class StateItem(object):
    """One quantity recorded from an iterative algorithm (interface of
    micmec/sampling/iterative.py:108-222).

    ``update(iterative)`` stores ``value = get_value(iterative)``; the first update also fixes
    ``shape`` and ``dtype`` (an ndarray's own, otherwise ``()`` and the Python type), which the
    trajectory writers use to lay out their datasets. ``iter_attrs`` yields ``(name, value)``
    pairs stored once beside the data. Subclasses provide ``get_value``.
    """

    def __init__(self, key):
        self.key = key
        self.shape = None
        self.dtype = None

    def get_value(self, iterative):
        raise NotImplementedError

    def update(self, iterative):
        value = self.get_value(iterative)
        self.value = value
        if self.shape is None:
            self.shape, self.dtype = _layout_of(value)

    def iter_attrs(self, iterative):
        return iter(())

    def copy(self):
        """A fresh item of the same kind and key, its shape not yet known."""
        twin = copy.copy(self)
        twin.shape = twin.dtype = None
        twin.__dict__.pop("value", None)
        return twin


def _layout_of(value):
    if isinstance(value, np.ndarray):
        return value.shape, value.dtype
    return (), type(value)


class AttributeStateItem(StateItem):
    """The iterative's attribute named ``key`` (None where it has none)."""

    def get_value(self, iterative):
        return getattr(iterative, self.key, None)


class ConsErrStateItem(StateItem):
    """The attribute ``key`` of the integrator's conserved-energy tracker."""

    def get_value(self, iterative):
        return getattr(iterative._cons_err_tracker, self.key, None)


def _fixed_item(name, key, getter, attrs=None, doc=None):
    """A StateItem subclass with a fixed key: ``getter(iterative)`` gives the value and
    ``attrs(iterative)`` the (name, value) attribute pairs."""

    def __init__(self):
        StateItem.__init__(self, key)

    body = {"__init__": __init__, "__doc__": doc,
            "get_value": lambda self, iterative: getter(iterative)}
    if attrs is not None:
        body["iter_attrs"] = lambda self, iterative: iter(attrs(iterative))
    return type(name, (StateItem,), body)


# key, value and attributes of the fixed items (iterative.py:147-192)
PosStateItem = _fixed_item(
    "PosStateItem", "pos", lambda it: it.mmf.system.pos,
    doc="Node positions of the iterative's system.")
TemperatureStateItem = _fixed_item(
    "TemperatureStateItem", "temp", lambda it: getattr(it, "temp", None),
    attrs=lambda it: [("ndof", it.ndof)],
    doc="Instantaneous temperature; the number of degrees of freedom is stored as attribute.")
VolumeStateItem = _fixed_item(
    "VolumeStateItem", "volume", lambda it: it.mmf.system.domain.volume,
    doc="Volume of the periodic domain.")
DomainStateItem = _fixed_item(
    "DomainStateItem", "domain", lambda it: it.mmf.system.domain.rvecs,
    doc="Domain vectors (rows).")
EPotContribStateItem = _fixed_item(
    "EPotContribStateItem", "epot_contribs",
    lambda it: np.array([part.energy for part in it.mmf.parts]),
    attrs=lambda it: [("epot_contrib_names",
                       np.array([part.name for part in it.mmf.parts], dtype="S"))],
    doc="Energy of every force part; their names are stored as attribute.")


class Hook(object):
    """A routine that an iterative algorithm calls at iterations start, start + step,
    start + 2 step, ... (interface of iterative.py:200-222)."""
    name = None
    kind = None
    method = None

    def __init__(self, start=0, step=1):
        self.start = start
        self.step = step

    def expects_call(self, counter):
        offset = counter - self.start
        return offset >= 0 and offset % self.step == 0

    def __call__(self, iterative):
        raise NotImplementedError
