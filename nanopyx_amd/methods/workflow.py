"""Workflow of analysis steps (src/nanopyx/methods/workflow.py:4-139): each step is (fn, args, kwargs);
"PREV_RETURN_VALUE_i_j" placeholders take the j-th return value of step i; the Agent picks the run type
unless ``_force_run_type`` is given, and is informed of the run's time afterwards."""
from __future__ import annotations

from ..__agent__ import Agent


class Workflow:
    def __init__(self, *args) -> None:
        self._methods = []
        self._return_values = []
        for arg in args:
            if isinstance(arg, tuple) and len(arg) == 3:
                self._methods.append((arg[0], arg[1], arg[2]))
            else:
                raise TypeError("Each arg must be a tuple of 3 items (fn, args, kwargs)")

    def _resolve(self, value):
        if isinstance(value, str) and "PREV_RETURN_VALUE" in value:
            idx = [int(i) for i in value.split("_") if i.isdigit()]
            return self._return_values[idx[0]][idx[1]]
        return value

    def run(self, _force_run_type=None):
        fn = None
        for fn, args, kwargs in self._methods:
            sane_args = [self._resolve(a) for a in args]
            kwargs = {k: self._resolve(v) for k, v in kwargs.items()}
            kwargs["run_type"] = Agent.get_run_type(fn, sane_args, kwargs)
            if _force_run_type:
                kwargs["run_type"] = _force_run_type
            rv = fn.run(*sane_args, **kwargs)
            self._return_values.append(rv if isinstance(rv, tuple) else (rv,))
            Agent._inform(fn)
        return self._return_values[-1], fn._last_runtype, fn._last_time

    def calculate(self, _force_run_type=None):
        output, _, _ = self.run(_force_run_type=_force_run_type)
        return output
