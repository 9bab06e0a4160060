"""Host-only plan of a call (bp_plan of include/beamform_gpu.h; no
counterpart in the reference): the device memory a call needs, whether it
fits a budget, and -- given the moveouts and weights -- the tile classes,
LDS windows, row padding and source order the engine would use.  Touches
no device, so admission decisions can be made and tested anywhere."""
import ctypes as ct

import numpy as np

from .core import PlanRequest, PlanResult, check, load_library

FORMATS = {"f32": 0, "bf16": 1, "f16": 2}
ORDERS = {"natural": 0, "similarity": 1, "auto": 2}


def plan(n_sources, n_stations, n_phases, n_samples, fmt="f32", reduce="max", hbm_budget=0,
         time_delays=None, weights_sources=None, order="auto"):
    """Returns the plan as a dict (fields of bp_plan_result)."""
    lib = load_library("gpu")
    req = PlanRequest()
    req.n_sources, req.n_stations = int(n_sources), int(n_stations)
    req.n_phases, req.n_samples = int(n_phases), int(n_samples)
    req.fmt = FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
    req.reduce = {"max": 0, "none": 1}[reduce] if isinstance(reduce, str) else int(reduce)
    req.hbm_budget = int(hbm_budget)
    req.order = ORDERS[order] if isinstance(order, str) else int(order)
    keep = []
    if time_delays is not None:
        tau = np.ascontiguousarray(time_delays, dtype=np.int32)
        beta = np.ascontiguousarray(weights_sources, dtype=np.float32)
        if tau.shape != (n_sources, n_stations, n_phases) or beta.shape != (n_sources, n_stations):
            raise ValueError("time_delays / weights_sources do not match the shape")
        keep = [tau, beta]
        req.time_delays, req.weights_sources = tau.ctypes.data, beta.ctypes.data
    res = PlanResult()
    check(lib.bp_plan(ct.byref(req), ct.byref(res)), "bp_plan", lib)
    del keep
    out = {}
    for name, _ in PlanResult._fields_:
        v = getattr(res, name)
        out[name] = list(v) if isinstance(v, ct.Array) else v
    out["fits"] = bool(out["fits"])
    return out
