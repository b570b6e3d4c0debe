"""MI355X-native deterministic / path propagation (witwin.channel surface).

Public entry points: :func:`witwin.channel.deterministic.solve`,
:func:`witwin.channel.path.solve` and
:func:`witwin.channel.propagation.consumer.evaluate`. All compute runs in
libwcmi.so (hand-written gfx950 HIP kernels) through a C ABI.
"""

from __future__ import annotations


def build_info() -> dict:
    from witwin.channel import _native

    lib = _native.load()
    return {
        "backend": "hip",
        "target": lib.wc_build_target().decode(),
        "abi_version": int(lib.wc_abi_version()),
        "library": str(_native.library_path()),
        # the reference's RayD / CUDA / OptiX backends are not used: every
        # solver runs on the HIP library above
        "uses_rayd_native": False,
        "uses_path_native": True,
        "uses_hip_native": True,
        "cuda_available": False,
        "optix_available": False,
    }


__all__ = ["build_info"]
