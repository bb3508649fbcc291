"""The reference's configuration files (include/rwnmr/rwnmr_config.h):
rwnmr.config, uct.config, cpmg.config, pfgse.config, multitau.config read
with the reference's line grammar (B/ConfigFiles/*_config.cpp)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import check, load


class ConfigFile:
    """Key/value store filled file by file (defaults first, then the user's)."""

    def __init__(self, *paths):
        h = ct.c_void_p()
        check(load().rwnmr_config_create(ct.byref(h)))
        self.handle = h
        for p in paths:
            self.read(p)

    def read(self, path):
        check(load().rwnmr_config_read(self.handle, str(path).encode()))
        return self

    def set(self, key, value):
        check(load().rwnmr_config_set(self.handle, key.encode(), str(value).encode()))
        return self

    def __contains__(self, key):
        return bool(load().rwnmr_config_has(self.handle, key.encode()))

    def string(self, key):
        buf = ct.create_string_buffer(4096)
        check(load().rwnmr_config_string(self.handle, key.encode(), buf, 4096))
        return buf.value.decode()

    def double(self, key):
        v = ct.c_double()
        check(load().rwnmr_config_double(self.handle, key.encode(), ct.byref(v)))
        return v.value

    def int(self, key):
        v = ct.c_int64()
        check(load().rwnmr_config_int(self.handle, key.encode(), ct.byref(v)))
        return v.value

    def bool(self, key):
        v = ct.c_int32()
        check(load().rwnmr_config_bool(self.handle, key.encode(), ct.byref(v)))
        return bool(v.value)

    def list(self, key, append=False):
        n = ct.c_uint32()
        check(load().rwnmr_config_list(self.handle, key.encode(), int(append), None, 0, ct.byref(n)))
        out = np.zeros(n.value)
        check(load().rwnmr_config_list(self.handle, key.encode(), int(append), ct.c_void_p(out.ctypes.data),
                                       n.value, ct.byref(n)))
        return out

    def close(self):
        if self.handle:
            load().rwnmr_config_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
