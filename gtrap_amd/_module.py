"""PyCUDA ``SourceModule``-shaped objects backed by libgtrap_hip.so.

The reference builds each kernel with ``SourceModule(...)`` and launches it with
``mod.get_function(name)(args..., block=, grid=, shared=)``
(gtrap/geebls.py:10-22, gtrap/gtls.py:11-27, gtrap/gfilter.py:8-89).  Here the
kernels are precompiled; ``get_function`` returns a callable with the same
argument order whose launch keywords keep their meaning (block[0] = nthread or
nw, grid = (nf|nt|cols, nq|rows)).  ``shared`` is accepted and ignored: LDS use
is decided by the library.
"""
from __future__ import annotations

from typing import Callable


class Function:
    def __init__(self, name: str, launcher: Callable):
        self.name = name
        self._launch = launcher

    def __call__(self, *args, block=(1, 1, 1), grid=(1, 1), shared: int = 0, stream=None,
                 **kw):
        block = tuple(int(b) for b in block) + (1,) * (3 - len(block))
        grid = tuple(int(g) for g in grid) + (1,) * (3 - len(grid))
        return self._launch(*args, block=block, grid=grid, shared=int(shared), stream=stream,
                            **kw)

    def __repr__(self) -> str:
        return f"<gtrap_amd Function {self.name}>"


class Module:
    def __init__(self, functions: dict[str, Callable]):
        self._functions = dict(functions)

    def get_function(self, name: str) -> Function:
        try:
            return Function(name, self._functions[name])
        except KeyError:
            raise KeyError(f"no kernel named {name!r}; available: {sorted(self._functions)}") \
                from None
