"""Host <-> device copies (mirror of melvin/DataTransferer.py:7-23)."""
from . import hip


class DataTransferer:
    def __init__(self, xp=hip):
        self._xp = hip

    def to_host(self, data):
        return hip.asnumpy(data)

    def from_host(self, data):
        return hip.asarray(data)
