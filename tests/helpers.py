"""Shared helpers for the parity tests."""
import numpy as np

import imex_amd as im

G = im.GC


def interior(a):
    return a[G:-G, G:-G]


def rel_l2(a, b):
    """Relative L2 distance of a from the reference b (0 when both vanish)."""
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0:
        return d
    return d / nb


def field_errors(u_test, u_ref):
    return [rel_l2(interior(u_test)[..., k], interior(u_ref)[..., k]) for k in range(u_ref.shape[-1])]


def make_case(name, n=None, m=None, amp=None, **over):
    p, kw = im.case_of(name)
    if n is not None:
        p.nx = n
        p.ny = m or n
    for k, v in over.items():
        setattr(p, k, v)
    a = kw.get("amp", 1.0) if amp is None else amp
    u, bx, by = im.init_case(p, kw["case"], a, kw.get("u0", 0.0), kw.get("v0", 0.0), kw.get("smooth", 0))
    return p, u, bx, by


def make_case_a(name, n=None, m=None, amp=None, **over):
    """make_case plus the case's vector potential (for the UNSTAGGER / POTENTIAL_A treatments)."""
    p, u, bx, by = make_case(name, n, m, amp, **over)
    _, kw = im.case_of(name)
    A = im.init_potential(p, kw["case"], kw.get("amp", 1.0) if amp is None else amp)
    return p, u, bx, by, A
