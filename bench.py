#!/usr/bin/env python3
"""Benchmark of the vring RK3 substep on one MI355X.

Workload: nr=256, ntheta=128, nz=512, r in [0,4],
z in [0,12], Re=3000, one Gaussian-core ring (R=1, z0=3, sigma=0.2, Gamma=1)
with seeded core-radius perturbations of amplitude 1e-3 on modes m=2..8
(phases from numpy PCG64(0)), dt at CFL 0.3, inputs resident in HBM.
A step is one full RK3 step (3 substeps: boundary terms, fused explicit
stencils, ADI sweeps, FFT/DCT/tridiagonal pressure solve, projection and
pressure update).  The Lagrangian particle step is timed separately on the
same field and reported as particles/s.

`value` is the mixed-precision path (FP32 stencils, sweeps and transforms;
FP64 for the m=0..2 pressure modes, the predictor, the projection and the
state) in Mcell-RK3-steps/s; the FP64 path, the ratio and the mixed path's
error against the FP64 path are reported beside it.  The roofline object is
the dominant kernel of the reported path, timed with HIP events on the
solver's stream over the timed region; the CPU baseline is the FP64 oracle
(a restatement of the reference algorithm) on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
NT, NR, NZ, RMAX, ZMAX, RE = 128, 256, 512, 4.0, 12.0, 3000.0
# algorithmic bytes per cell of the reference phases fused in each kernel
# (phase model: A hdnl 6 passes, B explicit RHS 18, C ADI 21, D divq 4,
#  E Poisson 10, F upsol 7, G uppres 18 -> 84 per substep, 252 per step)
PASSES_PER_STEP = 252


def kernel_alg_bytes_per_cell(name, prec):
    """Algorithmic bytes per cell of one launch of kernel `name` (the phase
    model's operands that the kernel's phases stream, each at its element
    size in the path: the state q, predictor dq, phi, p and the pressure
    gradient prg are FP64 in both paths; the work arrays ru, qrhs, the
    spectral array and qcap take the work type)."""
    w = 8 if prec == "fp64" else 4   # work arrays
    s = 8                            # FP64 operands
    if name == "explicit":
        # A: q1..3 in, H1..3 out; B per component: q, H, ru, prg in; qrhs, ru out
        return 3 * s + 3 * w + 3 * (s + w + w + s + w + w)
    if name == "project":
        # F: dq1..3 and phi in, q1..3 out; G per component: phi in, qcap out,
        # qcap in, prg in, prg out; then phi, p in, p out
        return (3 * s + s + 3 * s) + 3 * (s + w + w + s + s) + 3 * s
    if name == "line_r":
        # per launch, averaged over the three: q1 and q3 r sweeps rhs in/out;
        # q2's r sweep (after its z sweep) rhs and q2 in, dq2 out
        return (2 * w + 2 * w + (w + 2 * s)) / 3.0
    if name == "line_z":
        # q1 and q3 z sweeps rhs and q in, dq out; q2's z sweep rhs in/out
        return ((w + 2 * s) + (w + 2 * s) + 2 * w) / 3.0
    if name == "divfft":             # D: dq(3) in, rhs out; E: FFT in/out
        return 3 * s + w + w + w
    if name in ("dct_fwd", "dct_inv"):
        return 2 * w                 # spectral array in/out
    if name == "radial":
        return 3 * w                 # spectral array in/out, pivots in
    if name == "ifft":
        return w + s                 # spectral array in, phi out
    return None


def ring_state(nt=NT, nr=NR, nz=NZ):
    import vring_amd
    phases = np.random.Generator(np.random.PCG64(0)).uniform(0.0, 2.0 * np.pi, 7)
    q = vring_amd.ring_initial_condition(nt, nr, nz, RMAX, ZMAX, [(1.0, 3.0, 0.2, 1.0)], pert_amp=1e-3,
                                         pert_modes=(2, 8), phases=phases)
    dt = vring_amd.cfl_dt(nt, nr, nz, RMAX, ZMAX, *q, cfl=0.3, ujet=1.0)
    return q, dt


# the large single-GPU workload: (ntheta, nr, nz, rmax, zmax, Re)
LARGE = (256, 512, 1024, 5.0, 20.0, 7500.0)


def large_state():
    """Two coaxial Gaussian rings (R = 1 at z = 3 and R = 1.2 at z = 4:
    separation 1, sigma = 0.15, Gamma = 1) with 1e-3 core-radius
    perturbations on m = 2..8 (phases from numpy PCG64(0)) at 512x256x1024,
    r in [0,5], z in [0,20]; dt at CFL 0.3."""
    import vring_amd
    nt, nr, nz, rmax, zmax, _ = LARGE
    phases = np.random.Generator(np.random.PCG64(0)).uniform(0.0, 2.0 * np.pi, 7)
    q = vring_amd.ring_initial_condition(nt, nr, nz, rmax, zmax, [(1.0, 3.0, 0.15, 1.0), (1.2, 4.0, 0.15, 1.0)],
                                         pert_amp=1e-3, pert_modes=(2, 8), phases=phases)
    dt = vring_amd.cfl_dt(nt, nr, nz, rmax, zmax, *q, cfl=0.3, ujet=1.0)
    return q, dt


def run_large(steps, warmup, particles=16 * 1024 * 1024, tune_reps=3):
    """The large workload on one GPU: FP64 and mixed paths, K timed steps
    after W warm-up steps each; the mixed path carries `particles`
    trilinear-RK3 particles in the two rings' torus (stepped inside each RK3
    step) and is timed with and without them; the mixed-vs-FP64 error after
    the warm-up and timed steps."""
    import vring_amd
    nt, nr, nz, rmax, zmax, Re = LARGE
    q, dt = large_state()
    ncell = nt * nr * nz
    out = {"workload": f"nr={nr} ntheta={nt} nz={nz} r<={rmax} z<={zmax} Re={Re}, two rings", "dt": dt,
           "steps": steps, "warmup": warmup}
    fields = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(nt, nr, nz, rmax, zmax, Re, dt, precision=prec)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        tune = s.autotune(reps=tune_reps)
        s.step(warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        wall = time.perf_counter() - t0
        row = {"ms_per_step": round(wall / steps * 1e3, 3),
               "Mcell_RK3_steps_per_s": round(ncell * steps / wall / 1e6, 2), "autotune": tune}
        if prec == "mixed" and particles:
            s.seed_particles_torus(particles, R=1.1, z0=3.5, a=0.6, seed=0, scheme="trilinear_rk3")
            s.step(1)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(steps)
            wallp = time.perf_counter() - t0
            row["with_particles"] = {"particles": particles, "ms_per_step": round(wallp / steps * 1e3, 3),
                                     "particle_ms_per_step": round((wallp - wall) / steps * 1e3, 3),
                                     "particles_per_s": round(particles * steps / max(wallp - wall, 1e-9), 1)}
        fields[prec] = s
        out[prec] = row
    # the mixed path ran `steps + 1` more steps with particles: compare after
    # bringing the FP64 path to the same step
    extra = steps + 1 if particles else 0
    fields["fp64"].step(extra)
    f = {p: {n: s.download(n) for n in ("q1", "q2", "q3", "pres")} for p, s in fields.items()}
    vs = max(np.abs(f["fp64"][n]).max() for n in ("q1", "q2", "q3"))
    err = {n: float(np.abs(f["mixed"][n] - f["fp64"][n]).max() / vs) for n in ("q1", "q2", "q3")}
    err["pres"] = float(np.abs(f["mixed"]["pres"] - f["fp64"]["pres"]).max() / np.abs(f["fp64"]["pres"]).max())
    out["mixed_rel_err_vs_fp64"] = {"after_steps": warmup + steps + extra, "errors": err}
    out["mixed_over_fp64"] = round(out["mixed"]["Mcell_RK3_steps_per_s"] / out["fp64"]["Mcell_RK3_steps_per_s"], 3)
    for s in fields.values():
        s.close()
    return out


def run_path(prec, q, dt, steps, warmup, tune_reps=8):
    """Flow solver only: K timed RK3 steps after W warm-up steps."""
    import vring_amd
    s = vring_amd.Solver(NT, NR, NZ, RMAX, ZMAX, RE, dt, precision=prec)
    for n, a in zip(("q1", "q2", "q3"), q):
        s.upload(n, a)
    tuned = s.autotune(reps=tune_reps)
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)          # returns after the stream has drained
    t1 = time.perf_counter()
    wall = t1 - t0
    fields = {n: s.download(n) for n in ("q1", "q2", "q3", "pres")}
    # instrumented pass over the same number of steps: per-kernel HIP events
    s.enable_ktimer(True)
    s.step(steps)
    kt = s.kernel_times()
    s.enable_ktimer(False)
    s.close()
    return wall, fields, kt, tuned


def run_particles(prec, q, dt, n, steps, scheme="lsq_euler"):
    """Particle step on the workload's field: n particles uniform in the
    ring's torus, timed over `steps` calls of one RK3 step's particle work.
    lsq_euler: the reference's scheme (27-point least squares, Euler, 5 um
    droplets of batch 4 of part_tools.F90:262-269, nsubdt = 5, called at
    irk = 1 as loop.F90:67); trilinear_rk3: trilinear interpolation and the
    RK3 stages (one per substep)."""
    import vring_amd
    s = vring_amd.Solver(NT, NR, NZ, RMAX, ZMAX, RE, dt, precision=prec)
    for name, a in zip(("q1", "q2", "q3"), q):
        s.upload(name, a)
    nsub = 5 if scheme == "lsq_euler" else 1
    s.seed_particles_torus(n, R=1.0, z0=3.0, a=0.6, seed=0, nsub=nsub, scheme=scheme)
    stages = (1,) if scheme == "lsq_euler" else (1, 2, 3)
    for irk in stages:
        s.particles_step(irk)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for irk in stages:
            s.particles_step(irk)
    s.synchronize()
    t = time.perf_counter() - t0
    # the cell sort that upload and compaction run (vrh_particles_sort)
    ts = time.perf_counter()
    for _ in range(3):
        s.particles_sort(0)
    s.synchronize()
    sort_ms = (time.perf_counter() - ts) / 3 * 1e3
    s.close()
    nvar = 117 if scheme == "lsq_euler" else 6
    # state rows in and out (FP64); LSQ: the 81-value stencil per component
    # set; trilinear: the previous rates in and out (float rows in mixed)
    if scheme == "lsq_euler":
        per = (2 * nvar + 81) * 8
    else:
        per = 2 * nvar * 8 + 2 * nvar * (4 if prec == "mixed" else 8)
    return {"scheme": scheme, "particles": n, "substeps_per_step": nsub * len(stages),
            "ms_per_step": round(t / steps * 1e3, 3), "particles_per_s": round(n * steps / t, 1),
            "hbm_bytes_per_particle_substep": per,
            "hbm_GBps": round(per * n * nsub * len(stages) * steps / t / 1e9, 1),
            "order": "sorted by cell in device memory (upload and compaction sort)",
            "sort_ms": round(sort_ms, 3)}


def particle_errors(q, dt, n, scheme):
    """The mixed path's particle step against the FP64 one from the same
    field and particles: one RK3 step's particle work (LSQ-Euler: five
    sub-steps at irk = 1; trilinear RK3: its three stages), then the largest
    difference of the live particles' positions and of their velocities,
    each relative to the largest FP64 magnitude."""
    import vring_amd
    res = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(NT, NR, NZ, RMAX, ZMAX, RE, dt, precision=prec)
        for name, a in zip(("q1", "q2", "q3"), q):
            s.upload(name, a)
        nsub = 5 if scheme == "lsq_euler" else 1
        s.seed_particles_torus(n, R=1.0, z0=3.0, a=0.6, seed=0, nsub=nsub, scheme=scheme)
        for irk in ((1,) if scheme == "lsq_euler" else (1, 2, 3)):
            s.particles_step(irk)
        u, alive = s.particles_download(0, n)
        res[prec] = (u[:, :6].copy(), alive)
        s.close()
    (u64, a64), (umx, amx) = res["fp64"], res["mixed"]
    live = (a64 == 1) & (amx == 1)
    rel = lambda c: float(np.abs(umx[live][:, c] - u64[live][:, c]).max() / np.abs(u64[live][:, c]).max())
    return {"particles": n, "positions": rel(slice(0, 3)), "velocities": rel(slice(3, 6)),
            "alive_differs": int((a64 != amx).sum())}


def ensemble_specs(nrings, seed=0):
    """(nr, ntheta, nz, Re, sigma) of the batched-ensemble rings: shapes from
    {64, 96, 128, 192} x {32, 64, 128} x {128, 256, 512}, Re ~ U[500, 5000],
    core sigma ~ U[0.1, 0.3], drawn with numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nrings):
        nr, nt, nz = int(rng.choice([64, 96, 128, 192])), int(rng.choice([32, 64, 128])), int(rng.choice([128, 256, 512]))
        Re, sigma = float(rng.uniform(500, 5000)), float(rng.uniform(0.1, 0.3))
        out.append((nr, nt, nz, Re, sigma))
    return out


def ensemble_member(spec, particles, index):
    """One ensemble ring (mixed precision, dt at CFL 0.3) with `particles`
    trilinear-RK3 particles in its core's torus (seed = index)."""
    import vring_amd
    nr, nt, nz, Re, sigma = spec
    q = vring_amd.ring_initial_condition(nt, nr, nz, RMAX, ZMAX, [(1.0, 3.0, sigma, 1.0)])
    dt = vring_amd.cfl_dt(nt, nr, nz, RMAX, ZMAX, *q, cfl=0.3, ujet=1.0)
    s = vring_amd.Solver(nt, nr, nz, RMAX, ZMAX, Re, dt, precision="mixed")
    for n, a in zip(("q1", "q2", "q3"), q):
        s.upload(n, a)
    if particles:
        s.seed_particles_torus(particles, R=1.0, z0=3.0, a=3 * sigma, seed=index, nsub=1, scheme="trilinear_rk3")
    return s


def run_ensemble(nrings, steps, warmup, seed=0, particles=50000):
    """Batched ensemble: nrings independent rings with (nr, ntheta, nz) drawn
    from {64, 96, 128, 192} x {32, 64, 128} x {128, 256, 512}, Re ~ U[500, 5000],
    core sigma ~ U[0.1, 0.3] (numpy PCG64(seed)), mixed precision, `particles`
    trilinear-RK3 particles each, advanced by one graph launch per RK3 step.
    Each distinct grid shape is autotuned once (its first ring) and the
    others of that shape take the same launch shapes.  Returns aggregate
    Mcell-RK3-steps/s and the same rings stepped one after another with
    their own graphs."""
    import vring_amd
    members, shapes, cells, tuned = [], [], 0, {}
    t_tune = time.perf_counter()
    for i, spec in enumerate(ensemble_specs(nrings, seed)):
        s = ensemble_member(spec, particles, i)
        nr, nt, nz = spec[:3]
        if (nr, nt, nz) not in tuned:
            tuned[(nr, nt, nz)] = s.autotune(reps=2)
        else:
            c = tuned[(nr, nt, nz)]
            s.set_tuning(M=[c[k] for k in ("M_r1", "M_r3", "M_r2", "M_z1", "M_z3", "M_z2")],
                         fft_lines=c["fft_lines"], dct_lines=c["dct_lines"], explicit_form=c["explicit_form"])
        members.append(s)
        shapes.append((nr, nt, nz))
        cells += nr * nt * nz
    t_tune = time.perf_counter() - t_tune
    ens = vring_amd.Ensemble(members)
    ens.step(warmup)
    t0 = time.perf_counter()
    ens.step(steps)
    t_ens = time.perf_counter() - t0
    # all `steps` steps of every ring in one graph launch
    ens.set_batch(steps)
    ens.step(steps)
    t0 = time.perf_counter()
    ens.step(steps)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    for s in members:
        s.step(steps)
    t_seq = time.perf_counter() - t0
    ens.close()
    for s in members:
        s.close()
    return {"rings": nrings, "cells": cells, "steps": steps, "particles_per_ring": particles,
            "ms_per_step": round(t_ens / steps * 1e3, 3),
            "Mcell_RK3_steps_per_s": round(cells * steps / t_ens / 1e6, 2),
            "one_launch_per_run_Mcell_RK3_steps_per_s": round(cells * steps / t_batch / 1e6, 2),
            "one_by_one_Mcell_RK3_steps_per_s": round(cells * steps / t_seq / 1e6, 2),
            "autotuned_shapes": len(tuned), "setup_and_tuning_s": round(t_tune, 1),
            "shapes_nr_nt_nz": shapes}


def run_learned(nr, nz, nslots, train_samples, seed=0, pool=8):
    """Per-mode (r, z) Helmholtz on the tanh-stretched z grid (stretch 2),
    modes m = 0..nslots-1: train the network on `train_samples` (rhs, p)
    pairs the direct solver generates on the GPU, then solve ring-core right-hand
    sides to 1e-8 relative residual with plain PCG, BiCGSTAB, the learned
    flexible CG, and the direct solve."""
    import vring_amd
    hz = vring_amd.Helmholtz(nr, nz, nslots, RMAX, ZMAX, 2.0, pool=pool)
    hz.set_modes(np.arange(nslots, dtype=np.int32))
    hz.net_init(seed)
    t0 = time.perf_counter()
    loss = hz.train(train_samples, lr=1e-3, seed=seed)
    t_train = time.perf_counter() - t0
    hz.set_modes(np.arange(nslots, dtype=np.int32))
    f = hz.sample(seed + 99, 0)
    out = {"nr": nr, "nz": nz, "modes": f"0..{nslots - 1}", "network_grid": [nz // pool, nr // pool],
           "train_samples": train_samples,
           "train_s": round(t_train, 2), "loss_first": float(loss[:4].mean()), "loss_last": float(loss[-4:].mean())}
    pd = hz.direct(f)
    t0 = time.perf_counter()
    pd = hz.direct(f)
    out["direct_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
    for m in ("pcg", "bicgstab", "learned", "learned_fcg", "jacobi_warm_start"):
        w = None
        if m == "jacobi_warm_start":  # the learned method with the network switched off
            w = hz.net_get()
            hz.net_set(np.zeros_like(w))
        p, it, rel, ms = hz.solve(f, "learned" if w is not None else m, tol=1e-8,
                                  maxit=5000 if m == "learned_fcg" else 50000)
        if w is not None:
            hz.net_set(w)
        out[m] = {"iterations_total": int(it.sum()), "iterations_max": int(it.max()), "ms": round(ms, 3),
                  "max_true_relres": float(rel.max()),
                  "max_err_vs_direct": float(np.abs(p - pd).max() / np.abs(pd).max())}
    # the same network trained on the residuals of up to 256 plain PCG
    # iterations (the inputs a preconditioner meets), as the flexible CG's
    # preconditioner
    w_rhs = hz.net_get()
    hz.net_init(seed)
    t0 = time.perf_counter()
    lr_ = hz.train_residuals(train_samples, lr=1e-3, seed=seed, max_cg=256)
    hz.set_modes(np.arange(nslots, dtype=np.int32))
    p, it, rel, ms = hz.solve(f, "learned_fcg", tol=1e-8, maxit=5000)
    out["learned_fcg_residual_trained"] = {
        "train_s": round(time.perf_counter() - t0 - ms / 1e3, 2), "loss_first": float(lr_[:4].mean()),
        "loss_last": float(lr_[-4:].mean()), "iterations_total": int(it.sum()), "iterations_max": int(it.max()),
        "ms": round(ms, 3), "max_true_relres": float(rel.max()),
        "max_err_vs_direct": float(np.abs(p - pd).max() / np.abs(pd).max())}
    hz.net_set(w_rhs)
    # loose tolerances (inner solves that need only a few digits)
    for tol in (1e-2, 1e-4):
        row = {}
        for m in ("pcg", "learned", "jacobi_warm_start"):
            w = None
            if m == "jacobi_warm_start":
                w = hz.net_get()
                hz.net_set(np.zeros_like(w))
            p, it, rel, ms = hz.solve(f, "learned" if w is not None else m, tol=tol, maxit=50000)
            if w is not None:
                hz.net_set(w)
            row[m] = {"iterations_total": int(it.sum()), "ms": round(ms, 3), "max_true_relres": float(rel.max())}
        out[f"tol_{tol:g}"] = row
    hz.close()
    return out


def tanh_axial_mesh(nz, zmax, s):
    """Axial mesh with faces z_k = zmax (1 + tanh(s (k/nz - 1)) / tanh s),
    as vrh_set_axial_mesh takes it: centres, faces, dz/dzeta at centres and
    faces over k = 0..nz+1, mirrored ghost cells."""
    k = np.arange(nz + 1)
    zf = zmax * (1.0 + np.tanh(s * (k / nz - 1.0)) / np.tanh(s))
    zf[0], zf[-1] = 0.0, zmax
    zc = np.concatenate([[-zf[1]], zf])
    zm = np.zeros(nz + 2)
    zm[1:nz + 1] = 0.5 * (zf[1:] + zf[:-1])
    zm[0], zm[nz + 1] = -zm[1], 2.0 * zmax - zm[nz]
    dz = zmax / nz
    dmz = np.zeros(nz + 2)
    dmz[1:nz + 1] = (zf[1:] - zf[:-1]) / dz
    dmz[0], dmz[nz + 1] = dmz[1], dmz[nz]
    dcz = np.zeros(nz + 2)
    dcz[1:] = (zm[1:] - zm[:-1]) / dz
    dcz[0] = dcz[1]
    return zm, zc, dmz, dcz


def run_learned_step(train_samples, steps=2, warmup=1, stretch=2.0, mmin=3, tol=1e-8, seed=0, pool=8,
                     methods=("pcg", "learned", "learned_fcg"), max_cg=64):
    """The learned form inside the time step: the workload's grid with the
    tanh-stretched axial mesh (stretch 2, the per-mode Helmholtz grid), mixed
    precision, the pressure modes m >= mmin solved per mode to relative
    residual `tol` by plain PCG, PCG from the learned warm start, or the
    learned flexible CG (network trained on `train_samples` pairs, as the
    residuals of up to `max_cg` plain PCG iterations), against
    the direct eigen-transform solve of the same step.  Times are whole
    steps (eager launches: the axial mesh's z transforms are rocBLAS GEMMs);
    errors are against the direct mixed and FP64 runs after the same steps."""
    import vring_amd
    q, dt = ring_state()
    dt = dt / 8.0  # the first cells are 7x thinner than the uniform grid's
    mesh = tanh_axial_mesh(NZ, ZMAX, stretch)
    ncell = NT * NR * NZ
    nslots = NT - (2 * mmin - 1)
    hz = vring_amd.Helmholtz(NR, NZ, nslots, RMAX, ZMAX, stretch, pool=pool)
    hz.net_init(seed)
    t0 = time.perf_counter()
    loss = hz.train_residuals(train_samples, lr=1e-3, seed=seed, max_cg=max_cg)
    out = {"workload": f"nr={NR} ntheta={NT} nz={NZ} Re={RE:g}, tanh axial mesh stretch {stretch:g}",
           "modes_iterative": f"m >= {mmin} ({nslots} theta positions)", "tol": tol, "dt": dt,
           "train_samples": train_samples, "train_on": f"CG residuals (0..{max_cg} iterations)",
           "train_s": round(time.perf_counter() - t0, 2),
           "loss_first": float(loss[:4].mean()), "loss_last": float(loss[-4:].mean())}
    fields = {}
    runs = [("fp64_direct", "fp64", None), ("direct", "mixed", None)] + [(m, "mixed", m) for m in methods]
    for name, prec, method in runs:
        s = vring_amd.Solver(NT, NR, NZ, RMAX, ZMAX, RE, dt, precision=prec)
        s.set_axial_mesh(*mesh)
        if method:
            s.set_mode_solver(hz, method, mmin, tol, 20000)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.set_time(0.5)
        s.step(warmup)
        s.synchronize()
        it0, ns0 = s.mode_solver_stats()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        it1, ns1 = s.mode_solver_stats()
        fields[name] = {n: s.download(n) for n in ("q1", "q2", "q3", "pres")}
        row = {"ms_per_step": round(wall / steps * 1e3, 3),
               "Mcell_RK3_steps_per_s": round(ncell * steps / wall / 1e6, 2)}
        if method:
            row["krylov_iterations_per_solve"] = round((it1 - it0) / max(ns1 - ns0, 1), 1)
            row["iterations_per_slot_solve"] = round((it1 - it0) / max(ns1 - ns0, 1) / nslots, 1)
        if name != "fp64_direct":
            row["err_vs_mixed_direct"] = rel_errors(fields[name], fields["direct"]) if name != "direct" else None
            row["err_vs_fp64_direct"] = rel_errors(fields[name], fields["fp64_direct"])
        out[name] = row
        s.close()
    hz.close()
    if "learned" in out:
        out["learned_over_direct"] = round(out["direct"]["ms_per_step"] / out["learned"]["ms_per_step"], 4)
        if "pcg" in out:
            out["learned_over_pcg"] = round(out["pcg"]["ms_per_step"] / out["learned"]["ms_per_step"], 4)
    return out


def error_run(q, dt, steps=500, every=50):
    """Mixed against FP64 from the same initial field over `steps` steps
    (the workload's run length): per field, the largest error over the
    checkpoints every `every` steps, relative to the velocity scale for q1,
    q2, q3 and to max|p| for p, and the errors at the last step."""
    import vring_amd
    S = {}
    for prec in ("fp64", "mixed"):
        S[prec] = vring_amd.Solver(NT, NR, NZ, RMAX, ZMAX, RE, dt, precision=prec)
        for n, a in zip(("q1", "q2", "q3"), q):
            S[prec].upload(n, a)
    worst = {n: 0.0 for n in ("q1", "q2", "q3", "pres")}
    last = None
    for _ in range(every, steps + 1, every):
        for s in S.values():
            s.step(every)
        f = {p: {n: s.download(n) for n in worst} for p, s in S.items()}
        last = rel_errors(f["mixed"], f["fp64"])
        worst = {n: max(worst[n], last[n]) for n in worst}
    for s in S.values():
        s.close()
    return {"steps": steps, "checkpoint_every": every, "worst": worst, "final": last,
            "bar": 1e-6, "within_bar": max(worst.values()) < 1e-6}


def rel_errors(a, b):
    vs = max(np.abs(b[n]).max() for n in ("q1", "q2", "q3"))
    out = {n: float(np.abs(a[n] - b[n]).max() / vs) for n in ("q1", "q2", "q3")}
    out["pres"] = float(np.abs(a["pres"] - b["pres"]).max() / max(np.abs(b["pres"]).max(), 1e-300))
    return out


def roofline(kt, prec, ncell):
    """The dominant kernel (most time per step) of a path: algorithmic bytes
    per launch over its average launch time (HIP events on the solver's
    stream), and the measured HBM traffic per launch (rocprofv3 counters,
    profiles/traffic_256x128x512.json) over the same time."""
    best = max(kt.items(), key=lambda kv: kv[1][0])
    name, (ms, cnt) = best
    per_cell = kernel_alg_bytes_per_cell(name, prec)
    avg_ms = ms / cnt
    if per_cell is None:
        return {"kernel": name, "avg_ms": avg_ms, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None}
    ach = per_cell * ncell / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_256x128x512.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get(prec, {}).get(name)
    out = {"kernel": name, "avg_ms": round(avg_ms, 4), "bound": "hbm", "achieved": round(ach, 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 3), "traffic": traffic,
           "alg_bytes_per_cell": per_cell, "alg_bytes_per_launch": int(per_cell * ncell)}
    if traffic:
        out["counter_GBps"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
        out["frac_counter"] = round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)
    return out


def host_cpu():
    """CPU model and logical core count of the host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(q, dt):
    """FP64 oracle (liboracle.so, one host core) on one RK3 substep of the
    same workload at full size; Mcell-RK3-steps/s = cells / (3 x substep)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(NT, NR, NZ, RMAX, ZMAX, RE, dt)
    for n, a in zip(("q1", "q2", "q3"), q):
        o.field(n)[...] = a
    o.set_time(dt)
    t0 = time.perf_counter()
    o.substep(1)
    t = time.perf_counter() - t0
    o.close()
    model, ncpu = host_cpu()
    return {"value": round(NT * NR * NZ / (3.0 * t) / 1e6, 4), "unit": "Mcell-RK3-steps/s", "cores": 1,
            "kind": "port", "sample": f"one RK3 substep at 256x128x512 took {t:.1f} s (x3 per step)",
            "cpu_model": model, "host_logical_cores": ncpu}


def copy_calibration(mib, reps=3):
    """Plain device-to-device copy of `mib` MiB (torch elementwise copy
    kernel): reads and writes exactly mib MiB per launch.  Profiling it with
    the FETCH_SIZE / WRITE_SIZE counters gives the scale factor between the
    counters and real HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "reps": reps}))


KERNEL_KEYS = {  # rocprof kernel name prefix -> kernel timer name
    "k_explicit_cell": "explicit", "k_explicit_march": "explicit", "k_explicit_wave": "explicit",
    "k_explicit": "explicit", "k_lsq_fit": "particles", "k_lsq_step": "particles", "k_lsq_substep": "particles",
    "k_project": "project", "k_divfft": "divfft", "k_dctz": "dct", "k_radial": "radial", "k_ifft": "ifft",
    "k_hdnbc": "hdnbc", "k_part_lsq": "particles", "k_part_tri": "particles", "k_scalars": "scalars",
}


def _kernel_key(name):
    """('explicit', 'fp64') from a demangled rocprof kernel name."""
    base = name.split("(anonymous namespace)::", 1)[-1]
    fn = base.split("<")[0].split("(")[0]
    targs = base[len(fn) + 1:].split(">")[0] if "<" in base else ""
    prec = "fp64" if targs.startswith("double") else ("mixed" if targs.startswith("float") else "")
    if fn == "k_line":
        mode = int(targs.split(",")[2])
        key = "line_r" if mode in (0, 1, 2) else "line_z"
    else:
        key = KERNEL_KEYS.get(fn, fn)
    return key, prec


def summarize_traffic(pmc_root, out_path):
    """HBM bytes per launch of each kernel from rocprofv3 --pmc runs
    (pmc_<prec>_FETCH_SIZE, pmc_<prec>_WRITE_SIZE), each counter scaled by
    its reading on the 1 GiB plain copy (pmc_copy_*)."""
    import csv

    def rows(d):
        with open(os.path.join(pmc_root, d, "p_counter_collection.csv")) as f:
            return list(csv.DictReader(f))

    copy = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = [float(r["Counter_Value"]) for r in rows(f"pmc_copy_{c}") if "copyBuffer" in r["Kernel_Name"]]
        copy[c] = (1 << 30) / 1024.0 / (sum(vals) / len(vals))  # true KiB per counted KiB
    out = {"unit": "bytes per launch", "scale": {"FETCH_SIZE": copy["FETCH_SIZE"], "WRITE_SIZE": copy["WRITE_SIZE"]},
           "calibration": "1 GiB device copy: true bytes / counter bytes"}
    for prec in ("fp64", "mixed"):
        acc = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            for r in rows(f"pmc_{prec}_{c}"):
                key, kp = _kernel_key(r["Kernel_Name"])
                if kp != ("fp64" if prec == "fp64" else "mixed"):
                    continue
                acc.setdefault(key, {}).setdefault(c, []).append(float(r["Counter_Value"]) * copy[c] * 1024.0)
        out[prec] = {k: round(sum(v["FETCH_SIZE"]) / len(v["FETCH_SIZE"]) + sum(v["WRITE_SIZE"]) / len(v["WRITE_SIZE"]))
                     for k, v in acc.items() if "FETCH_SIZE" in v and "WRITE_SIZE" in v}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out))


def summarize_stats(csv_path, out_path, command, title="workload 256x128x512 ring (bench.py)", raw=False):
    """Text summary of a rocprofv3 --kernel-trace --stats kernel_stats.csv:
    per kernel (timer name, or the kernel's own name with raw=True, and
    precision) calls and average duration."""
    import csv
    with open(csv_path) as f:
        rows = list(csv.DictReader(f))
    agg = {}
    for r in rows:
        key, prec = _kernel_key(r["Name"])
        if raw:
            base = r["Name"].split("(anonymous namespace)::", 1)[-1]
            key = base.split("<")[0].split("(")[0].replace("void ", "")
            if key.startswith("Cijk_"):  # rocBLAS/Tensile GEMM: keep the operand layout part
                key = key.split("_MT")[0] + " (rocBLAS GEMM)"
        a = agg.setdefault((key, prec or "-"), [0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    tot = sum(v[1] for v in agg.values())
    note = (["Kernels by their own names; `type` is the work type of the instantiation."] if raw else
            ["Kernel names are mapped to the solver's kernel-timer names; `type` is the",
             "work type of the instantiation (in the mixed path the FP64 low-mode",
             "transforms and radial solves appear as fp64).  Calls include the",
             "autotuner's timing launches."])
    lines = [f"# rocprofv3 kernel statistics, {title}", "", f"command: `{command}`", ""] + note + [
             "", "| kernel | type | calls | avg us | total ms | % |", "|---|---|---|---|---|---|"]
    for (key, prec), (calls, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| {key} | {prec} | {calls} | {ns / calls / 1e3:.1f} | {ns / 1e6:.2f} | {100 * ns / tot:.1f} |")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=1000000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--error-steps", type=int, default=500,
                    help="steps of the mixed-vs-FP64 error run (0: skip)")
    ap.add_argument("--autotune-reps", type=int, default=8)
    ap.add_argument("--only", choices=("fp64", "mixed", "particles"), default=None,
                    help="run one path only and print its timing (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="RINGS",
                    help="run only the batched-ensemble benchmark with RINGS rings")
    ap.add_argument("--large", action="store_true",
                    help="run only the large workload (512x256x1024, two rings, 16M particles)")
    ap.add_argument("--learned", type=int, default=0, metavar="SAMPLES",
                    help="run only the learned Helmholtz benchmark, training on SAMPLES pairs")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 64-ring ensemble and the large workload in the default run")
    ap.add_argument("--learned-train", type=int, default=8192, metavar="SAMPLES",
                    help="training pairs of the learned-in-the-step measurement of the default run (0: skip it)")
    ap.add_argument("--learned-step", type=int, default=0, metavar="SAMPLES",
                    help="run only the learned-in-the-step benchmark (tanh axial mesh), training on SAMPLES pairs")
    ap.add_argument("--summarize-traffic", metavar="PMC_DIR", default=None,
                    help="write profiles/traffic_256x128x512.json from rocprofv3 --pmc output under PMC_DIR")
    ap.add_argument("--stats-title", default="workload 256x128x512 ring (bench.py)",
                    help="title of the --summarize-stats summary")
    ap.add_argument("--stats-raw-names", action="store_true",
                    help="--summarize-stats lists kernels by their own names")
    ap.add_argument("--summarize-stats", nargs=3, metavar=("CSV", "OUT", "COMMAND"), default=None,
                    help="write a text summary of a rocprofv3 kernel_stats.csv")
    args = ap.parse_args()
    if args.summarize_stats:
        summarize_stats(*args.summarize_stats, title=args.stats_title, raw=args.stats_raw_names)
        return
    if args.summarize_traffic:
        summarize_traffic(args.summarize_traffic, os.path.join(ROOT, "profiles", "traffic_256x128x512.json"))
        return
    if args.gpus != 1:
        print("bench.py: vring_amd runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.ensemble:
        print(json.dumps({"ensemble": run_ensemble(args.ensemble, args.steps, args.warmup)}))
        return
    if args.large:
        print(json.dumps({"large": run_large(args.steps, args.warmup, args.particles * 16)}))
        return
    if args.learned:
        print(json.dumps({"learned": run_learned(256, 512, 64, args.learned)}))
        return
    if args.learned_step:
        print(json.dumps({"learned_step": run_learned_step(args.learned_step, args.steps, args.warmup)}))
        return
    q, dt = ring_state()
    ncell = NT * NR * NZ
    if args.only == "particles":
        rows = {f"{prec}_{sch}": run_particles(prec, q, dt, args.particles, 5, sch)
                for prec in ("fp64", "mixed") for sch in ("lsq_euler", "trilinear_rk3")}
        rows["mixed_rel_err_vs_fp64"] = {sch: particle_errors(q, dt, args.particles, sch)
                                         for sch in ("lsq_euler", "trilinear_rk3")}
        print(json.dumps(rows))
        return
    if args.only:
        wall, _, kt, tune = run_path(args.only, q, dt, args.steps, args.warmup, args.autotune_reps)
        print(json.dumps({"path": args.only, "ms_per_step": round(wall / args.steps * 1e3, 4),
                          "Mcell_RK3_steps_per_s": round(ncell * args.steps / wall / 1e6, 2),
                          "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items()},
                          "autotune": tune}))
        return
    wall64, f64, kt64, tune64 = run_path("fp64", q, dt, args.steps, args.warmup, args.autotune_reps)
    wallmx, fmx, ktmx, tunemx = run_path("mixed", q, dt, args.steps, args.warmup, args.autotune_reps)
    parts = None
    if args.particles:
        ps = max(2, args.steps // 4)
        parts = {f"{prec}_{sch}": run_particles(prec, q, dt, args.particles, ps, sch)
                 for prec in ("fp64", "mixed") for sch in ("lsq_euler", "trilinear_rk3")}
        parts["mixed_rel_err_vs_fp64"] = {sch: particle_errors(q, dt, args.particles, sch)
                                          for sch in ("lsq_euler", "trilinear_rk3")}
    v64 = ncell * args.steps / wall64 / 1e6
    vmx = ncell * args.steps / wallmx / 1e6
    err = error_run(q, dt, args.error_steps) if args.error_steps else {"timed_run": rel_errors(fmx, f64)}
    lstep = run_learned_step(args.learned_train, 1, 1) if args.learned_train else None
    # the other workloads' single-GPU measurements, beside the headline
    extras = {} if args.no_extras else {"ensemble_64": run_ensemble(64, 10, 2), "large": run_large(10, 2)}
    rf = roofline(ktmx, "mixed", ncell)
    rf64 = roofline(kt64, "fp64", ncell)
    step_frac64 = PASSES_PER_STEP * 8 * ncell * (args.steps / wall64) / (HBM_PEAK_GBS * 1e9)
    # the mixed path's phase-model bytes per step: the FP64 operands of the
    # 84 passes at 8 B, the work arrays at 4 B
    step_bytes_mx = 3 * ncell * sum(kernel_alg_bytes_per_cell(k, "mixed") * n for k, n in (
        ("explicit", 1), ("line_r", 3), ("line_z", 3), ("divfft", 1), ("dct_fwd", 1), ("radial", 1),
        ("dct_inv", 1), ("ifft", 1), ("project", 1)))
    step_frac_mx = step_bytes_mx * (args.steps / wallmx) / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": "Mcell-RK3-steps/s at 256x128x512, mixed+learned vs plain FP64 @1e-6 err",
        "value": round(vmx, 2),
        "unit": "Mcell-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wallmx / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)",
        "form": ("mixed precision: FP32 stencils, ADI sweeps, theta FFT, z DCT and radial solves of m >= 3; "
                 "FP64 state, predictor, m = 0..2 pressure modes and the projection with the pressure "
                 "gradient.  No learned solve in the timed step: on this uniform z grid the pressure solve is "
                 "the direct FFT + DCT + tridiagonal one; the learned form runs on the tanh-stretched axial "
                 "mesh and is measured beside it as learned_in_step (a loss)"),
        "data": "synthetic: seeded Gaussian ring with m=2..8 core perturbations",
        "config": {"workload": "nr=256 ntheta=128 nz=512 Re=3000 ring, 1 GPU", "nr": NR, "ntheta": NT,
                   "nz": NZ, "dt": dt, "particles": args.particles},
        "fp64": {"value": round(v64, 2), "ms_per_step": round(wall64 / args.steps * 1e3, 4),
                 "step_hbm_frac": round(step_frac64, 3), "roofline": rf64},
        "mixed_over_fp64": round(vmx / v64, 3),
        "step_hbm_frac": round(step_frac_mx, 3),
        "mixed_rel_err_vs_fp64": err,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in ktmx.items()},
        "fp64_kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt64.items()},
        "autotune": {"fp64": tune64, "mixed": tunemx},
        "particles": parts,
        "learned_in_step": lstep,
        **extras,
        "roofline": rf,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(q, dt)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
