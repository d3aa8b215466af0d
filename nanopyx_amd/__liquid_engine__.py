"""Minimal Liquid Engine base for the eSRRF operators.

NanoPyx's operators subclass ``LiquidEngine`` from the external ``liquid_engine`` package
(src/nanopyx/__liquid_engine__.py:1), which is not part of the reference repository. This base keeps the
observed contract (the reference's own call sites: src/nanopyx/core/transform/_le_esrrf.pyx:26-37,
src/nanopyx/methods/workflow.py:104-120): run types are discovered from ``_run_<name>`` methods,
``run(..., run_type=None)`` dispatches through ``_run``, ``benchmark(...)`` returns
``[(seconds, run_type_name, output), ...]`` fastest first, and each run sets ``_last_runtype`` and
``_last_time``. The only run type here is the gfx950 one, named ``HIP_gfx950``.
"""
from __future__ import annotations

import time
from collections import defaultdict

HIP_RUN_TYPE = "HIP_gfx950"


class LiquidEngine:
    _designation = "LiquidEngine"

    def __init__(self, clear_benchmarks: bool = False, testing: bool = False, verbose: bool = True):
        self._verbose = verbose
        self._testing = testing
        self._last_runtype = None
        self._last_time = None
        self._history = defaultdict(list)
        self._run_types = {}
        for attr in dir(type(self)):
            if attr.startswith("_run_") and callable(getattr(self, attr)):
                self._run_types[self._name_of_run_type(attr[len("_run_"):])] = getattr(self, attr)

    @staticmethod
    def _name_of_run_type(suffix: str) -> str:
        return HIP_RUN_TYPE if suffix == "hip" else suffix.capitalize()

    def _default_run_type(self) -> str:
        timed = {k: min(v) for k, v in self._history.items() if v and k in self._run_types}
        if timed:
            return min(timed, key=timed.get)
        return next(iter(self._run_types))

    def _run(self, *args, run_type=None, **kwargs):
        if run_type is None:
            run_type = self._default_run_type()
        if run_type not in self._run_types:
            raise ValueError(f"{self._designation}: unknown run type {run_type!r}; "
                             f"available: {sorted(self._run_types)}")
        t0 = time.perf_counter()
        out = self._run_types[run_type](*args, **kwargs)
        dt = time.perf_counter() - t0
        self._last_runtype = run_type
        self._last_time = dt
        self._history[run_type].append(dt)
        if self._verbose:
            print(f"Agent: {self._designation} using {run_type} ran in {dt} seconds")
        return out

    def benchmark(self, *args, **kwargs):
        results = []
        for name in self._run_types:
            out = self._run(*args, run_type=name, **kwargs)
            results.append((self._last_time, name, out))
        return sorted(results, key=lambda r: r[0])
