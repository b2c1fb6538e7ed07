"""CPU oracle of the learned per-mode Helmholtz solve (TEST INFRASTRUCTURE ONLY).

The reference has no learned component and no iterative pressure solve on
its substep path; its Poisson solve is the direct FFT/DCT/tridiagonal one
(poisson_solve.F90:140-179, 270-323).  This module restates, in numpy, what
libvrh's Helmholtz subsystem computes, so that tests can pin it:

* the per-Fourier-mode (r, z) pressure operator of the cylindrical Poisson
  equation on a uniform r grid and a tanh-stretched z grid, in symmetric
  finite-volume form  S_m p = V f  (fluxes as in poisson.F90:368-388, the
  exact theta eigenvalue m^2/r^2 of 'wvnex', poisson_solve.F90:286-289,
  zero-flux axis, wall and z ends);
* its direct solution (sparse LU) with the m = 0 gauge fixed by a zero
  volume-weighted mean;
* plain CG and BiCGSTAB in the textbook form (the algorithmic form of the
  reference's VR2D/cg.F90 and VR2D/bicgstab.F90, which are not on the solver
  path);
* the 4-layer 3x3 convolutional network (NHWC, bf16 storage of activations,
  weights and back-propagated gradients, fp32 accumulation) with its
  backward pass and the energy-norm loss of the preconditioned estimate
  D^-1 V f + scale N(f) (Jacobi plus the learned correction).

Only tests/ may import this module.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ----------------------------------------------------------------------------
# mesh and operator
# ----------------------------------------------------------------------------


def mesh(nr, nz, rmax, zmax, stretch):
    """Uniform r cells; z faces z_k = L (1 + tanh(s (k/nz - 1)) / tanh s),
    k = 0..nz (clustered towards z = 0, the inlet); stretch = 0 gives a uniform grid."""
    dr = rmax / nr
    rm = (np.arange(nr) + 0.5) * dr
    rface = np.arange(nr + 1) * dr
    k = np.arange(nz + 1)
    if stretch > 0:
        zf = zmax * (1.0 + np.tanh(stretch * (k / nz - 1.0)) / np.tanh(stretch))
    else:
        zf = zmax * k / nz
    zm = 0.5 * (zf[1:] + zf[:-1])
    dz = zf[1:] - zf[:-1]
    return dict(nr=nr, nz=nz, dr=dr, rm=rm, rface=rface, zf=zf, zm=zm, dz=dz)


def coefficients(g):
    """Face coefficients of S: ar[k, j] on r faces j = 1..nr-1 (between
    cells j-1, j), az[k, j] on z faces k = 1..nz-1, the m^2 weight w and the
    cell volume V (all arrays (nz, nr) or the face subsets)."""
    dr, rm, rf, dz, zm = g["dr"], g["rm"], g["rface"], g["dz"], g["zm"]
    ar = dz[:, None] * (rf[None, 1:-1] / dr)            # (nz, nr-1)
    az = (rm[None, :] * dr) / (zm[1:] - zm[:-1])[:, None]  # (nz-1, nr)
    w = dz[:, None] * (dr / rm[None, :])                # (nz, nr)
    V = dz[:, None] * (rm[None, :] * dr)
    return ar, az, w, V


def apply(g, m, x):
    """y = S_m x for x of shape (nz, nr) or (B, nz, nr) with per-sample m."""
    ar, az, w, _ = coefficients(g)
    x = np.asarray(x, np.float64)
    m = np.asarray(m, np.float64)
    mm = (m * m)[..., None, None] if m.ndim else m * m
    y = mm * w * x
    fr = ar * (x[..., :, 1:] - x[..., :, :-1])
    y[..., :, 1:] += fr
    y[..., :, :-1] -= fr
    fz = az * (x[..., 1:, :] - x[..., :-1, :])
    y[..., 1:, :] += fz
    y[..., :-1, :] -= fz
    return y


def diag(g, m):
    ar, az, w, _ = coefficients(g)
    d = m * m * w.copy()
    d[:, 1:] += ar
    d[:, :-1] += ar
    d[1:, :] += az
    d[:-1, :] += az
    return d


def matrix(g, m):
    nz, nr = g["nz"], g["nr"]
    n = nz * nr
    idx = np.arange(n).reshape(nz, nr)
    rows, cs, vals = [], [], []
    ar, az, w, _ = coefficients(g)
    d = diag(g, m)
    rows.append(idx.ravel()); cs.append(idx.ravel()); vals.append(d.ravel())
    rows += [idx[:, 1:].ravel(), idx[:, :-1].ravel()]
    cs += [idx[:, :-1].ravel(), idx[:, 1:].ravel()]
    vals += [-ar.ravel(), -ar.ravel()]
    rows += [idx[1:, :].ravel(), idx[:-1, :].ravel()]
    cs += [idx[:-1, :].ravel(), idx[1:, :].ravel()]
    vals += [-az.ravel(), -az.ravel()]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cs))), shape=(n, n))


def rhs(g, m, f):
    """b = V f; for m = 0 the volume-weighted mean of f is removed first
    (the solvability condition of the pure-Neumann problem)."""
    _, _, _, V = coefficients(g)
    f = np.asarray(f, np.float64)
    if m == 0:
        f = f - (V * f).sum() / V.sum()
    return V * f


def gauge(g, m, p):
    if m != 0:
        return p
    _, _, _, V = coefficients(g)
    return p - (V * p).sum() / V.sum()


def direct(g, m, f):
    """Exact solution of S_m p = V f (zero weighted mean for m = 0)."""
    b = rhs(g, m, f).ravel()
    A = matrix(g, m).tolil()
    if m == 0:  # pin one cell, then fix the gauge
        A[0, :] = 0.0
        A[0, 0] = 1.0
        b = b.copy()
        b[0] = 0.0
    p = spla.spsolve(A.tocsc(), b).reshape(g["nz"], g["nr"])
    return gauge(g, m, p)


def cg(g, m, f, tol, maxit, x0=None, jacobi=True):
    """Textbook preconditioned CG (Jacobi or none) on S_m p = V f."""
    b = rhs(g, m, f)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    dinv = 1.0 / diag(g, m) if jacobi else np.ones_like(b)
    r = b - apply(g, m, x)
    z = dinv * r
    p = z.copy()
    rz = (r * z).sum()
    nb = np.sqrt((b * b).sum())
    for it in range(1, maxit + 1):
        q = apply(g, m, p)
        a = rz / (p * q).sum()
        x += a * p
        r -= a * q
        if np.sqrt((r * r).sum()) <= tol * nb:
            return gauge(g, m, x), it
        z = dinv * r
        rz2 = (r * z).sum()
        p = z + (rz2 / rz) * p
        rz = rz2
    return gauge(g, m, x), maxit


def bicgstab(g, m, f, tol, maxit, x0=None):
    """Textbook Jacobi-preconditioned BiCGSTAB on S_m p = V f."""
    b = rhs(g, m, f)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    dinv = 1.0 / diag(g, m)
    r = b - apply(g, m, x)
    rh = r.copy()
    rho = alpha = om = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    nb = np.sqrt((b * b).sum())
    for it in range(1, maxit + 1):
        rho2 = (rh * r).sum()
        beta = (rho2 / rho) * (alpha / om)
        p = r + beta * (p - om * v)
        ph = dinv * p
        v = apply(g, m, ph)
        alpha = rho2 / (rh * v).sum()
        s = r - alpha * v
        sh = dinv * s
        t = apply(g, m, sh)
        om = (t * s).sum() / (t * t).sum()
        x += alpha * ph + om * sh
        r = s - om * t
        rho = rho2
        if np.sqrt((r * r).sum()) <= tol * nb:
            return gauge(g, m, x), it
    return gauge(g, m, x), maxit


# ----------------------------------------------------------------------------
# network
# ----------------------------------------------------------------------------

CIN0 = 16          # input channels as stored (4 features + 12 zero pads)
NFEAT = 4
WIDTH = 32
DIL = (1, 2, 4, 8)
LAYERS = ((CIN0, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, 16))  # last: channel 0 is the output


def bf16(x):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def features(g, m, f):
    """Input channels (nz, nr, 16): f / rms(f), m dr / (m dr + rm), dz / dr,
    rm / rmax, then zeros.  Returns (x, scale) with scale = rms(f) R^2
    (R = nr dr), the factor that turns the network output into a pressure
    correction."""
    nz, nr = g["nz"], g["nr"]
    s = float(np.sqrt(np.mean(np.asarray(f, np.float64) ** 2)))
    s = s if s > 0 else 1.0
    x = np.zeros((nz, nr, CIN0), np.float32)
    x[..., 0] = f / s
    x[..., 1] = (m * g["dr"]) / (m * g["dr"] + g["rm"][None, :])
    x[..., 2] = (g["dz"] / g["dr"])[:, None]
    x[..., 3] = (g["rm"] / (g["dr"] * nr))[None, :]
    return bf16(x), s * (g["dr"] * nr) ** 2


def init_params(seed):
    """He-style uniform init (fp32 master weights); W[l] (9, cout, cin), b[l] (cout)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ws, bs = [], []
    for (ci, co) in LAYERS:
        fan = 9 * (NFEAT if ci == CIN0 else ci)
        lim = np.sqrt(6.0 / fan)
        W = rng.uniform(-lim, lim, (9, co, ci)).astype(np.float32)
        if ci == CIN0:
            W[:, :, NFEAT:] = 0.0
        Ws.append(W)
        bs.append(np.zeros(co, np.float32))
    Ws[-1][:, 1:, :] = 0.0
    Ws[-1] *= 0.1
    return Ws, bs


def _shift(x, dk, dj):
    """out[k, j] = x[k + dk, j + dj] with zero padding (x: (..., H, W, C))."""
    out = np.zeros_like(x)
    H, W = x.shape[-3], x.shape[-2]
    ks, ke = max(0, -dk), min(H, H - dk)
    js, je = max(0, -dj), min(W, W - dj)
    if ke > ks and je > js:
        out[..., ks:ke, js:je, :] = x[..., ks + dk:ke + dk, js + dj:je + dj, :]
    return out


TAPS = [(dk, dj) for dk in (-1, 0, 1) for dj in (-1, 0, 1)]  # tap t = 3 (dk+1) + (dj+1)


def conv(x, W, b, dil):
    """y[p, co] = b[co] + sum_t sum_ci W[t, co, ci] x[p + dil*tap_t, ci] (float64 sums)."""
    y = np.zeros(x.shape[:-1] + (W.shape[1],))
    x64 = x.astype(np.float64)
    for t, (dk, dj) in enumerate(TAPS):
        y += _shift(x64, dk * dil, dj * dil) @ W[t].astype(np.float64).T
    return y + b


def _id(x):
    return x


def forward(params, x, rounding=True):
    """Forward pass with the library's rounding points; returns the output
    channel (float32) and the stored activations (bf16 values).  With
    rounding=False everything stays float64 (for finite-difference checks)."""
    Ws, bs = params
    rd = bf16 if rounding else _id
    cast = (lambda a: a.astype(np.float32)) if rounding else _id
    acts = [x]
    h = x
    for l, (W, b) in enumerate(zip(Ws, bs)):
        y = cast(conv(h, rd(W), b, DIL[l]))
        if l < 3:
            h = rd(np.maximum(y, 0.0))
            acts.append(h)
        else:
            out = y[..., 0]
    return out, acts


def backward(params, acts, gout, rounding=True):
    """Gradients of sum(gout * out) with the library's rounding points:
    back-propagated activation gradients are stored as bf16."""
    Ws, bs = params
    rd = bf16 if rounding else _id
    gW, gb = [None] * 4, [None] * 4
    g = np.zeros(gout.shape + (16,), np.float32 if rounding else np.float64)
    g[..., 0] = gout
    g = rd(g)
    for l in (3, 2, 1, 0):
        W, dil, a = rd(Ws[l]), DIL[l], acts[l]
        g64, a64 = g.astype(np.float64), a.astype(np.float64)
        gw = np.zeros(W.shape)
        for t, (dk, dj) in enumerate(TAPS):
            xs = _shift(a64, dk * dil, dj * dil)
            gw[t] = g64.reshape(-1, g64.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
        gW[l] = gw.astype(np.float32) if rounding else gw
        gb[l] = g64.reshape(-1, g.shape[-1]).sum(0)
        if rounding:
            gb[l] = gb[l].astype(np.float32)
        if l == 0:
            break
        gx = np.zeros(a.shape)
        for t, (dk, dj) in enumerate(TAPS):
            gx += _shift(g64 @ W[t].astype(np.float64), -dk * dil, -dj * dil)
        gx = (gx.astype(np.float32) if rounding else gx) * (a > 0)
        g = rd(gx)
    return gW, gb


def prolong_matrix(n, pool):
    """(n, n // pool) bilinear prolongation in index space from block centres
    to cells, constant beyond the outermost centres: cell i sits at
    u = (i + 1/2) / pool - 1/2 in node units."""
    nn = n // pool
    P = np.zeros((n, nn))
    for i in range(n):
        u = (i + 0.5) / pool - 0.5
        if u <= 0:
            P[i, 0] = 1.0
        elif u >= nn - 1:
            P[i, nn - 1] = 1.0
        else:
            i0 = int(np.floor(u))
            P[i, i0] += 1.0 - (u - i0)
            P[i, i0 + 1] += u - i0
    return P


def prolong(out, pool):
    """Network-grid field (nz / pool, nr / pool) to the cells (nz, nr)."""
    nzn, nrn = out.shape
    return prolong_matrix(nzn * pool, pool) @ out.astype(np.float64) @ prolong_matrix(nrn * pool, pool).T


def galerkin(g, m, r, c):
    """Energy-optimal combination z = beta D^-1 r + gamma c approximating
    S^-1 r: (beta, gamma) solve [[j.Sj, j.Sc], [j.Sc, c.Sc]] (beta, gamma) =
    (j.r, c.r) with j = D^-1 r; a nearly dependent pair keeps beta j alone."""
    j = r / diag(g, m)
    Sj, Sc = apply(g, m, j), apply(g, m, c)
    a11, a12, a22 = (j * Sj).sum(), (j * Sc).sum(), (c * Sc).sum()
    r1, r2 = (j * r).sum(), (c * r).sum()
    det = a11 * a22 - a12 * a12
    beta, gam = (r1 / a11 if a11 > 0 else 0.0), 0.0
    if a11 > 0 and a22 > 0 and det > 1e-10 * a11 * a22:
        beta, gam = (r1 * a22 - r2 * a12) / det, (a11 * r2 - a12 * r1) / det
    return beta * j + gam * c, beta, gam


def energy_loss_grad(g, m, out, p, f, scale):
    """Energy-norm loss of one sample's Galerkin estimate, with b = V f,
    c = scale * out and e = beta D^-1 b + gamma c - p for the optimal
    (beta, gamma):  L = e^T S e / (p^T b);  dL/d out = 2 gamma scale S e /
    (p^T b) (the coefficients are stationary, so their variation drops out)."""
    b = rhs(g, m, f)
    z, beta, gam = galerkin(g, m, b, scale * out.astype(np.float64))
    e = z - p
    Se = apply(g, m, e)
    den = (p * b).sum()
    return (e * Se).sum() / den, (2.0 * gam * scale / den) * Se
