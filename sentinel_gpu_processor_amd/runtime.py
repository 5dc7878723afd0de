"""Process-wide library context used by the drop-in modules.

One context per process and device (S1IFG_DEVICE, default 0), bound to
torch's current stream on that device.
"""

import os
import threading

_lock = threading.Lock()
_ctx = {}


def device():
    return f"cuda:{int(os.environ.get('S1IFG_DEVICE', '0'))}"


def context():
    from . import _native as N
    dev = int(os.environ.get("S1IFG_DEVICE", "0"))
    with _lock:
        c = _ctx.get(dev)
        if c is None:
            c = N.Context(dev)
            _ctx[dev] = c
        return c
