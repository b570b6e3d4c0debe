"""Component vocabulary and the configuration rules shared by the solvers.

Names, sets and error messages follow the reference contract
(witwin/channel/components.py:9-243) so callers see the same validation.
"""

from __future__ import annotations

from collections.abc import Iterable

VALID_COMPONENTS = frozenset({"los", "reflection", "diffraction", "transmission", "scattering"})
DEFAULT_COMPONENTS = frozenset({"los", "reflection", "diffraction"})
BOUNCE_COMPONENTS = frozenset({"reflection", "diffraction", "transmission", "scattering"})
AD_MODES = frozenset({"none", "jvp", "vjp"})
DEPTH_CAPPED_COMPONENTS = frozenset({"reflection", "transmission"})
# components this build evaluates; scattering (Kirchhoff rough surfaces) is outside the hot path
SUPPORTED_COMPONENTS = frozenset({"los", "reflection", "diffraction", "transmission"})
MAX_DEPTH = 5


def validated_components(components: Iterable[str], *, error_message: str) -> frozenset[str]:
    normalized = frozenset(components)
    if not normalized or not normalized.issubset(VALID_COMPONENTS):
        raise ValueError(error_message.format(valid=sorted(VALID_COMPONENTS)))
    return normalized


def validate_max_depth(max_depth: int) -> None:
    if max_depth < 0:
        raise ValueError("max_depth must be non-negative")


def component_max_depth(components: Iterable[str], *, chain_depth: int, single_bounce_depth: int) -> dict[str, int]:
    requested = frozenset(components)
    return {
        "los": 0 if "los" in requested else -1,
        "reflection": chain_depth if "reflection" in requested else -1,
        "diffraction": single_bounce_depth if "diffraction" in requested else -1,
        "transmission": chain_depth if "transmission" in requested else -1,
        "scattering": single_bounce_depth if "scattering" in requested else -1,
    }


def component_bits(components: Iterable[str]) -> int:
    """Bit mask of the native kernels: los 1, reflection 2, diffraction 4, transmission 8."""
    bits = 0
    for name, bit in (("los", 1), ("reflection", 2), ("diffraction", 4), ("transmission", 8)):
        if name in components:
            bits |= bit
    return bits
