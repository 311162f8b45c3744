"""Block registry: every public callable is registered under a stable block id.

Same contract as the reference registry (src/gpufluid/blocks/__init__.py:45-150):
``block(id, desc)`` records (id, description, qualname, module, source file and
line) and tags the callable with ``__block_id__``; ``BlockError`` carries the
id of the block that failed; ids look like ``S2.6.1``.
"""
from __future__ import annotations

import inspect
import logging
import re
from dataclasses import dataclass, field
from typing import Callable, Dict, List

_ID_RE = re.compile(r"^[A-Z][0-9]+(\.[0-9A-Z]+){1,3}$")


@dataclass
class BlockInfo:
    block_id: str
    description: str
    qualname: str
    module: str
    source_file: str
    source_line: int
    layer: str = field(init=False)

    def __post_init__(self):
        m = re.match(r"^([A-Z][0-9]+)", self.block_id)
        self.layer = m.group(1) if m else "?"


_REGISTRY: Dict[str, List[BlockInfo]] = {}


def block(block_id: str, description: str = "") -> Callable:
    """Decorator registering ``fn`` under ``block_id``; a second registration
    of the same qualname under one id is dropped with a warning."""
    if not _ID_RE.match(block_id):
        raise ValueError(f"bad block_id {block_id!r}; expected like 'S2.6.1'")

    def wrap(fn):
        try:
            src = inspect.getsourcefile(fn) or "?"
            _, line = inspect.getsourcelines(fn)
        except (TypeError, OSError):
            src, line = "?", 0
        info = BlockInfo(block_id, description or (fn.__doc__ or "").strip(),
                         getattr(fn, "__qualname__", getattr(fn, "__name__", "?")),
                         getattr(fn, "__module__", "?"), src, line)
        bucket = _REGISTRY.setdefault(block_id, [])
        if any(b.qualname == info.qualname for b in bucket):
            logging.getLogger("gpufluid_amd.blocks").warning(
                "blocks: duplicate qualname %r under %r dropped", info.qualname, block_id)
        else:
            bucket.append(info)
        try:
            fn.__block_id__ = block_id
        except (AttributeError, TypeError):
            pass
        return fn

    return wrap


def get_registry() -> Dict[str, List[BlockInfo]]:
    return {k: list(v) for k, v in _REGISTRY.items()}


def find(block_id: str) -> List[BlockInfo]:
    return list(_REGISTRY.get(block_id, []))


def by_layer(layer: str) -> List[BlockInfo]:
    items = [b for bucket in _REGISTRY.values() for b in bucket if b.layer == layer]

    def key(b):
        out = []
        for part in b.block_id[len(layer):].lstrip(".").split("."):
            out.append((0, int(part)) if part.isdigit() else (1, part))
        return tuple(out)

    return sorted(items, key=key)


class BlockError(RuntimeError):
    """Error tagged with the block id that raised it: ``[S2.6.1] message``."""

    def __init__(self, block_id: str, message: str):
        self.block_id = block_id
        self.message = message
        super().__init__(f"[{block_id}] {message}")


def assert_block(condition: bool, block_id: str, message: str) -> None:
    if not condition:
        raise BlockError(block_id, message)


__all__ = ["BlockInfo", "BlockError", "block", "get_registry", "find", "by_layer",
           "assert_block"]
