"""Physical constants and the phase / amplitude convention of every result.

The values and convention strings are part of the public contract shared with
the reference (witwin/channel/constants.py:10-90): a field is the phasor
exp(-j k d) under exp(+j 2 pi f t) time dependence, free-space amplitude
wavelength / (4 pi d) scaled by sqrt(tx_power) for ``path_field``.
"""

from __future__ import annotations

import math

C0 = 299792458.0
MU0 = 4.0e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0 = MU0 * C0

PHASOR = "exp(-j*k*d)"
TIME_DEPENDENCE = "exp(+j*2*pi*f*t)"
FREE_SPACE_AMPLITUDE = "sqrt(tx_power)*wavelength/(4*pi*distance)"
UNIT_EXCITATION_FREE_SPACE_AMPLITUDE = "wavelength/(4*pi*distance)"
POLARIZATION = "world_cartesian_complex3_then_receiver_projection"

PHASE_CONVENTION = {
    "phasor": PHASOR,
    "time_dependence": TIME_DEPENDENCE,
    "free_space_amplitude": FREE_SPACE_AMPLITUDE,
    "polarization": POLARIZATION,
}
UNIT_EXCITATION_PHASE_CONVENTION = {**PHASE_CONVENTION, "free_space_amplitude": UNIT_EXCITATION_FREE_SPACE_AMPLITUDE}

NARROWBAND_FREQUENCY_OFFSET_LAW = "H(f_ref+df) = C(f_ref)*exp(-j*2*pi*df*delay_s)"

__all__ = [
    "C0", "EPS0", "ETA0", "FREE_SPACE_AMPLITUDE", "MU0", "NARROWBAND_FREQUENCY_OFFSET_LAW", "PHASE_CONVENTION",
    "PHASOR", "POLARIZATION", "TIME_DEPENDENCE", "UNIT_EXCITATION_FREE_SPACE_AMPLITUDE",
    "UNIT_EXCITATION_PHASE_CONVENTION",
]
