"""Run-type Agent of the Workflow (src/nanopyx/methods/workflow.py:104,118; the reference's Agent lives in
the external ``liquid_engine`` package): picks the fastest run type seen so far for an operator and records
the time of each finished run."""
from __future__ import annotations

from collections import defaultdict


class _Agent:
    def __init__(self):
        self._times = defaultdict(lambda: defaultdict(list))

    def get_run_type(self, fn, args, kwargs):
        hist = self._times.get(fn._designation)
        if hist:
            best = {k: min(v) for k, v in hist.items() if v and k in fn._run_types}
            if best:
                return min(best, key=best.get)
        return fn._default_run_type()

    def _inform(self, fn):
        if fn._last_runtype is not None and fn._last_time is not None:
            self._times[fn._designation][fn._last_runtype].append(fn._last_time)


Agent = _Agent()
