"""Base class of every module that runs on a backend (the reference's
``utils/parallelbase.py:8-11`` ``BBHxParallelModule``)."""

from __future__ import annotations


class BBHxParallelModule:
    """Resolves ``force_backend`` (a name or a backend instance) once and
    exposes it as ``self.backend``; ``self.xp`` is the backend's array module."""

    def __init__(self, force_backend=None):
        from .. import get_backend

        if force_backend is None or isinstance(force_backend, str):
            self._backend = get_backend(force_backend or "hip")
        else:
            self._backend = force_backend

    @property
    def backend(self):
        return self._backend

    @property
    def xp(self):
        return self._backend.xp

    @property
    def device(self):
        return self._backend.device

    @classmethod
    def supported_backends(cls) -> list:
        return ["bbhx_hip"]
