"""Helpers shared by the GPU parity tests: build matching oracle and GPU
states and compare them."""
import numpy as np

from oracle_lib import Oracle

import vring_amd

STATE = ["q1", "q2", "q3", "dq1", "dq2", "dq3", "ru1", "ru2", "ru3", "pres"]


def make_pair(nt, nr, nz, rmax=4.0, zmax=12.0, Re=1000.0, dt=0.01, slip=0.0, theta="exact_m2",
              precision="fp64"):
    o = Oracle(nt, nr, nz, rmax, zmax, Re, dt, slip, wvnex=1 if theta == "exact_m2" else 0)
    s = vring_amd.Solver(nt, nr, nz, rmax, zmax, Re, dt, slip, theta_eigen=theta, precision=precision)
    return o, s


def seed_random(o, s, seed):
    rng = np.random.default_rng(seed)
    nt, nr, nz = o.nt, o.nr, o.nz
    for n in STATE:
        a = rng.standard_normal((nt, nr + 3, nz + 3))
        o.field(n)[...] = a
        s.upload(n, a)
    prg = rng.standard_normal((nt, nr, nz, 3))
    o.field("prg")[...] = prg
    s.upload("prg", prg)
    # outlet planes: the oracle stores csort as (nt, 1:nr+1), the GPU as (nt, 0:nr+2);
    # dq3no is (nt, 0:nr+2) on both sides
    cs = rng.standard_normal((nt, nr + 1))
    o.field("csort")[...] = cs
    plane = np.zeros((nt, nr + 3))
    plane[:, 1 : nr + 2] = cs
    s.upload("csort", plane)
    d3 = rng.standard_normal((nt, nr + 3))
    o.field("dq3no")[...] = d3
    s.upload("dq3no", d3)

def seed_fields(o, s, q1, q2, q3):
    for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
        o.field(n)[...] = a
        s.upload(n, a)


def relerr(a, b, scale=None):
    if scale is None:
        scale = np.abs(b).max()
    if scale == 0.0:
        return np.abs(a).max()
    return np.abs(a - b).max() / scale


def compare(o, s, names, tol, scale=None):
    """Max error relative to the oracle field's max magnitude (or `scale`)."""
    errs = {}
    for n in names:
        if n == "prg":
            errs[n] = relerr(s.download("prg"), o.field("prg"), scale)
        elif n == "dph":
            g = s.download("dph")[:, 1 : o.nr + 1, 1 : o.nz + 2]
            errs[n] = relerr(g, o.field("dph")[:, 1 : o.nr + 1, 1 : o.nz + 2], scale)
        else:
            errs[n] = relerr(s.download(n), o.field(n), scale)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, (bad, errs)
    return errs
