"""MI355X-native YarnBall stepper.

``yarnball_amd.Sim`` mirrors the reference's ``YarnBall::Sim``
(KittenEngine/YarnBall/YarnBall.h:116-206) on top of the C ABI of
``libyarnball_hip.so`` (include/yarnball_hip.h).  The C++ drop-in for the same
surface is include/YarnBall/YarnBall.h.
"""
from ._lib import (  # noqa: F401
    FLAG_FIX_ORIENTATION,
    FLAG_HAS_NEXT,
    FLAG_HAS_NEXT_ORIENTATION,
    FLAG_HAS_PREV,
    MAX_COLLISIONS_PER_SEGMENT,
    VERTEX_DTYPE,
    Params,
    YarnBallError,
)
from .sim import Sim  # noqa: F401
