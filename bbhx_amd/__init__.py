"""bbhx_amd: an MI355X (gfx950) backend for BBHx's frequency-domain
template-to-likelihood path (PhenomHM -> LISA TDI response -> cubic spline ->
direct / heterodyned inner products).

The public classes mirror the reference package's: ``BBHWaveformFD``,
``TemplateInterpFD`` (``waveformbuild``), ``PhenomHMAmpPhase``
(``waveforms.phenomhm``), ``LISATDIResponse`` (``response.fastfdresponse``),
``CubicSplineInterpolant`` (``utils.interpolate``), ``Likelihood`` and
``HeterodynedLikelihood`` (``likelihood``). Backends are resolved by name with
``get_backend`` like the reference's registry (``__init__.py:51-72``); the one
backend here is ``"bbhx_hip"`` (aliases ``"hip"``, ``"gpu"``, ``"cuda"``,
``"rocm"``).
"""

from __future__ import annotations

__version__ = "0.1.0"

from .cutils import BackendUnavailableException, BBHxHipBackend

_BACKENDS = {"bbhx_hip": BBHxHipBackend}
_ALIASES = {"hip": "bbhx_hip", "gpu": "bbhx_hip", "cuda": "bbhx_hip", "rocm": "bbhx_hip",
            "cuda11x": "bbhx_hip", "cuda12x": "bbhx_hip"}
_INSTANCES = {}


def _canonical(name: str) -> str:
    if name is None:
        return "bbhx_hip"
    name = name.lower()
    if name.startswith("bbhx_"):
        short = name[len("bbhx_"):]
        return _ALIASES.get(short, name)
    return _ALIASES.get(name, "bbhx_" + name)


def get_backend(backend: str = "hip"):
    """Return the backend instance registered under ``backend`` (with or
    without the ``bbhx_`` prefix). Raises BackendUnavailableException when it
    cannot be used here (unknown name, no GPU, library not built)."""
    key = _canonical(backend)
    if key not in _BACKENDS:
        raise BackendUnavailableException(f"backend '{backend}' is not available; known: {sorted(_BACKENDS)}")
    if key not in _INSTANCES:
        _INSTANCES[key] = _BACKENDS[key]()
    return _INSTANCES[key]


def has_backend(backend: str = "hip") -> bool:
    try:
        get_backend(backend)
        return True
    except BackendUnavailableException:
        return False


def get_first_backend(backend=None):
    """First usable backend from a name or a list of names."""
    names = [backend] if isinstance(backend, str) or backend is None else list(backend)
    for n in names:
        if has_backend(n):
            return get_backend(n)
    raise BackendUnavailableException(f"none of {names} is available")


__all__ = ["get_backend", "has_backend", "get_first_backend", "BackendUnavailableException", "__version__"]
