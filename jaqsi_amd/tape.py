"""Thread-local operation tapes.

Mirrors jaqsi/tape.py:10-138: a per-thread stack of tapes, ``recording()``
pushes a fresh list, ``active_tape()`` returns the innermost one, and
``shift_and_append``/``copy_to_tape`` replay a side tape onto shifted wires.
"""

from __future__ import annotations

import copy
import threading
from contextlib import contextmanager
from typing import Callable, Iterator, List, Optional

_local = threading.local()


def _tape_stack() -> List[list]:
    if not hasattr(_local, "stack"):
        _local.stack = []
    return _local.stack


class RecordPerSample(Exception):
    """Raised while a circuit is recorded once for a whole batch (under
    vmap) when it reaches a gate that must be recorded per sample (a pulse
    gate, whose propagator is solved by a native launch)."""


@contextmanager
def vectorized_recording() -> Iterator[None]:
    """Mark the recording inside as one recording for a whole batch."""
    prev = getattr(_local, "vectorized", False)
    _local.vectorized = True
    try:
        yield
    finally:
        _local.vectorized = prev


def require_per_sample_recording() -> None:
    """Called by gates that cannot be recorded once for a batch."""
    if getattr(_local, "vectorized", False):
        raise RecordPerSample()


def active_tape() -> Optional[list]:
    """The innermost recording tape of this thread, or None (tape.py:24-34)."""
    stack = _tape_stack()
    return stack[-1] if stack else None


@contextmanager
def recording() -> Iterator[list]:
    """Record every Operation created inside the block (tape.py:37-55)."""
    stack = _tape_stack()
    tape: list = []
    stack.append(tape)
    try:
        yield tape
    finally:
        stack.pop()


def shift_and_append(tape_ops: list, offset: int) -> None:
    """Append shallow copies of *tape_ops* with wires shifted by *offset*
    to the active tape (tape.py:92-115)."""
    current = active_tape()
    if current is None:
        return
    for o in tape_ops:
        shifted = copy.copy(o)
        shifted._wires = [w + offset for w in o.wires]
        current.append(shifted)


def copy_to_tape(fn: Callable[[], None], offset: int) -> None:
    """Record *fn* on a side tape and replay it shifted (tape.py:118-138)."""
    with recording() as side:
        fn()
    shift_and_append(side, offset)
