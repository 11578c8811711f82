"""Cadence callback (mirror of melvin/Ticker.py:1-35): fires once the loop
counter (or simulated time) passes its next mark, then moves the mark by
the cadence; so a cadence-10 loop ticker fires at loops 1, 11, 21, ..."""


class Ticker:
    def __init__(self, cadence, fn, dump_name="", is_loop_counter=True):
        self._cadence = cadence
        self._fn = fn
        self._counter = 0
        self.times_fired = 0
        self._dump_name = dump_name
        self._loop = is_loop_counter

    def tick(self, t, loop_counter):
        mark = loop_counter if self._loop else t
        if self._counter < mark:
            self._fn(self)
            self._counter += self._cadence
            self.times_fired += 1

    def dump(self):
        return {"counter": self._counter, "times_fired": self.times_fired}

    def restore(self, data):
        self._counter = data["counter"]
        self.times_fired = data["times_fired"]

    def get_name(self):
        return self._dump_name
