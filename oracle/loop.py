"""Bounded while-loop semantics of utils/loop.py:21-115 (oracle code).

That reference file is Copyright 2021 Google LLC under the Apache License,
Version 2.0; this restates its loop schedule.

Only *when* cond and body run matters for the solver, so the loops are plain
Python with the same call pattern as the jitted scan forms.
"""
import numpy as np


def geometric_schedule(maxiter, cond_exp_base):
    """Steps between condition checks for cond_exp_base > 1 (utils/loop.py:85-98)."""
    num = np.ceil(np.log((cond_exp_base - 1) * maxiter - 1) / np.log(cond_exp_base)) + 1
    steps = (cond_exp_base ** np.arange(num)).astype(int)
    mask = np.cumsum(steps) >= maxiter
    steps = np.delete(steps, np.where(mask)[0])
    return np.append(steps, maxiter - np.sum(steps)).astype(int)


def while_loop(cond_fun, body_fun, init_val, maxiter, cond_exp_base=1.0):
    """cond(init); then while cond holds: run body (a whole schedule chunk of
    bodies when cond_exp_base > 1) and re-evaluate cond.  At most maxiter
    bodies (utils/loop.py:21-51)."""
    assert 1.0 <= cond_exp_base <= 2.0
    val = init_val
    cond = cond_fun(val)
    if cond_exp_base > 1.0:
        for steps in geometric_schedule(maxiter, cond_exp_base):
            if not cond:
                break
            for _ in range(int(steps)):
                val = body_fun(val)
            cond = cond_fun(val)
        return val
    for _ in range(maxiter):
        if not cond:
            break
        val = body_fun(val)
        cond = cond_fun(val)
    return val
