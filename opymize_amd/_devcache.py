"""Per-device, per-dtype copies of an operator's constant arrays."""
import numpy as np
import torch

from opymize_amd import _hip


class DeviceConstants(object):
    """Uploads host constants once per (name, dtype, device) and keeps them."""

    def __init__(self):
        self._store = {}

    def get(self, name, host, dtype):
        dev = torch.cuda.current_device()
        key = (name, dtype, dev)
        t = self._store.get(key)
        if t is None:
            t = _hip.to_device(np.asarray(host), dtype)
            self._store[key] = t
        return t

    def get_u8(self, name, host):
        dev = torch.cuda.current_device()
        key = (name, torch.uint8, dev)
        t = self._store.get(key)
        if t is None:
            _hip.require_device()
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(host, dtype=np.uint8))).to("cuda")
            self._store[key] = t
        return t

    def get_i32(self, name, host):
        dev = torch.cuda.current_device()
        key = (name, torch.int32, dev)
        t = self._store.get(key)
        if t is None:
            _hip.require_device()
            t = torch.as_tensor(np.ascontiguousarray(np.asarray(host, dtype=np.int32))).to("cuda")
            self._store[key] = t
        return t

    def drop(self, name=None):
        if name is None:
            self._store.clear()
        else:
            for k in [k for k in self._store if k[0] == name]:
                del self._store[k]
