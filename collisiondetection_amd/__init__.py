"""MI355X-native continuous collision detection for swept triangle meshes.

A from-scratch HIP (gfx950) engine for the swept-primitive path of evouga/collisiondetection:
KDOPBroadPhase -> CTCDNarrowPhase -> CTCD primitives, behind the C ABI of include/ccd_amd.h
(libccd_amd.so).  ``collisiondetection_amd.api`` holds the Python front end.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "api"]
