"""TEST INFRASTRUCTURE ONLY: FP64 NumPy restatements of the reference's
reconstruction loops (the callers of Ax/Atb/minimizeTV), driven by the FP64
oracle operators.  Used to check tigre_amd.algorithms on small problems.

Followed files (Python/tigre/):
  utilities/filtering.py:15-87          ramp/Shepp-Logan/... filtering
  algorithms/single_pass_algorithms.py:12-224   FDK, FBP
  utilities/order_subsets.py:31-96      ordered subsets
  algorithms/iterative_recon_alg.py:214-401     W, V, OS-SART update
  algorithms/krylov_subspace_algorithms.py:22-97 CGLS
  algorithms/pocs_algorithms.py:105-169 ASD-POCS, :434-646 PCSD
  utilities/im_3d_denoise.py:7-28       im3ddenoise (normalised ROF denoise)
  algorithms/art_family_algorithms.py:65-153  SART-TV, OS-SART-TV
  algorithms/ista_algorithms.py:113-209 FISTA, ISTA
  algorithms/krylov_subspace_algorithms.py:101-188 LSQR, :190-297 hybrid LSQR, :299-439 LSMR,
                                        :570-695 AB-GMRES, BA-GMRES
  algorithms/statistical_algorithms.py:11-55     MLEM
  algorithms/krylov_subspace_algorithms.py:441-567 IRN-TV-CGLS, :698-860 hybrid fLSQR-TV
"""
from __future__ import annotations

import copy

import numpy as np
from scipy.fft import fft, ifft

from . import atb, aw_pocs_tv, ax, pocs_tv, tvdenoise


def prep(geo, angles):
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    g.cast_to_single()
    return g


def sub_geometry(geo, idx):
    """Geometry restricted to the views idx (iterative_recon_alg.py:335-341)."""
    g = copy.deepcopy(geo)
    for a in ("DSD", "DSO", "COR"):
        setattr(g, a, np.asarray(getattr(geo, a))[idx])
    for a in ("offOrigin", "offDetector", "rotDetector"):
        setattr(g, a, np.asarray(getattr(geo, a))[idx])
    return g


# -- FDK ---------------------------------------------------------------------------
def ramp_filter(geo, filter_name=None):
    """Frequency response used by filtering (filtering.py:18-23, :56-87)."""
    n = int(geo.nDetector[1])
    i = 1
    while 2 ** i <= 2 * n:
        i += 1
    filt_len = max(64, 2 ** i)
    nn = np.arange(-filt_len / 2, filt_len / 2)
    h = np.zeros(nn.shape, dtype=np.float32)
    h[int(filt_len / 2)] = 0.25
    odd = nn % 2 == 1
    h[odd] = -1 / (np.pi * nn[odd]) ** 2
    f = np.abs(np.fft.fft(h)) * 2
    filt = f[: int(filt_len / 2 + 1)]
    w = 2 * np.pi * np.arange(len(filt)) / filt_len
    d = 1
    if filter_name == "shepp_logan":
        filt[1:] *= np.sin(w[1:] / (2 * d)) / (w[1:] / (2 * d))
    elif filter_name == "cosine":
        filt[1:] *= np.cos(w[1:] / (2 * d))
    elif filter_name == "hamming":
        filt[1:] *= 0.54 + 0.46 * np.cos(w[1:] / d)
    elif filter_name == "hann":
        filt[1:] *= (1 + np.cos(w[1:] / d)) / 2
    elif filter_name not in ("ram_lak", None):
        raise ValueError("filter not recognized: " + str(filter_name))
    filt[w > np.pi * d] = 0
    return np.hstack((filt, filt[1:-1][::-1])).astype(np.float32), filt_len


def wang_pad(proj, geo):
    """Offset-detector zero padding (single_pass_algorithms.py:78-111)."""
    g = copy.deepcopy(geo)
    g.offDetector = np.array(g.offDetector, dtype=np.float64)
    g.nDetector = np.array(g.nDetector)
    g.sDetector = np.array(g.sDetector, dtype=np.float64)
    off = float(g.offDetector[1])
    dU = float(geo.dDetector[1])
    width = int(2 * off / dU)
    g.offDetector[1] = off - width / 2 * dU
    g.nDetector[1] = abs(width) + int(geo.nDetector[1])
    g.sDetector[1] = g.nDetector[1] * dU
    theta = (float(geo.sDetector[1]) / 2 - abs(off)) * np.sign(off)
    z = np.zeros((proj.shape[0], proj.shape[1], abs(width)))
    out = np.concatenate((z, proj), axis=2) if off > 0 else np.concatenate((proj, z), axis=2)
    return out, g, theta


def wang_weight(geo, theta):
    """Wang redundancy weight over u (single_pass_algorithms.py:113-163)."""
    DSO, DSD = float(np.ravel(geo.DSO)[0]), float(np.ravel(geo.DSD)[0])
    offset = float(geo.offDetector[1]) + DSD / DSO * float(np.ravel(getattr(geo, "COR", 0.0))[0])
    n = int(geo.nDetector[1])
    w = np.ones(n)
    lim = abs(theta * DSO / DSD)
    for k in range(n):
        t = ((k - n / 2 + 0.5) * float(geo.dDetector[1]) + abs(offset)) * DSO / DSD
        if abs(t) <= lim:
            w[k] = 0.5 * (np.sin(np.pi / 2 * np.arctan(t / DSO) / np.arctan(lim / DSO)) + 1)
        if t < -lim:
            w[k] = 0.0
    return w[::-1] if theta < 0 else w


def fdk(proj, geo, angles, filter_name=None, dowang=True):
    if dowang and np.any(np.asarray(geo.offDetector)):
        proj, geo, theta = wang_pad(np.asarray(proj, dtype=np.float64), geo)
        proj = proj * (2 * wang_weight(geo, theta))
    g = prep(geo, angles)
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    dV, dU = float(g.dDetector[0]), float(g.dDetector[1])
    DSD0, DSO0 = float(np.asarray(g.DSD)[0]), float(np.asarray(g.DSO)[0])
    xv = (np.arange(nU) - nU / 2 + 0.5) * dU
    yv = (np.arange(nV) - nV / 2 + 0.5) * dV
    xx, yy = np.meshgrid(xv, yv)
    w = DSD0 / np.sqrt(DSD0 ** 2 + xx ** 2 + yy ** 2)
    p = np.asarray(proj, dtype=np.float64) * w
    filt, L = ramp_filter(g, filter_name)
    pad = (L - nU) // 2
    scale = (DSD0 / DSO0) * (2 * np.pi / len(g.angles)) / (4 * dU)
    buf = np.zeros((p.shape[0], nV, L))
    buf[:, :, pad:pad + nU] = p
    out = np.real(ifft(fft(buf, axis=2) * filt.astype(np.float64), axis=2))[:, :, pad:pad + nU] * scale
    return atb(out.astype(np.float32), g, "FDK"), out


def fbp(proj, geo, angles, filter_name=None):
    """single_pass_algorithms.py:208-224 (parallel beam): the projections
    ramp-filtered as in FDK but without the cosine weight, backprojected and
    scaled by DSO / DSD (in float32) of the first view."""
    g = prep(geo, angles)
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    dU = float(g.dDetector[1])
    DSD0, DSO0 = float(np.asarray(g.DSD)[0]), float(np.asarray(g.DSO)[0])
    filt, L = ramp_filter(g, filter_name)
    pad = (L - nU) // 2
    scale = (DSD0 / DSO0) * (2 * np.pi / len(g.angles)) / (4 * dU)
    buf = np.zeros((len(g.angles), nV, L))
    buf[:, :, pad:pad + nU] = np.asarray(proj, dtype=np.float64)
    out = np.real(ifft(fft(buf, axis=2) * filt.astype(np.float64), axis=2))[:, :, pad:pad + nU] * scale
    return atb(out.astype(np.float32), g, "matched") * float(np.float32(DSO0 / DSD0))


# -- OS-SART -----------------------------------------------------------------------
def ordered_blocks(n, blocksize):
    nb = (n + blocksize - 1) // blocksize
    return [np.arange(i * blocksize, min(n, (i + 1) * blocksize)) for i in range(nb)]


def weights_w(geo, angles):
    """W = 1 / Ax(ones(2^3)) over an enlarged volume (iterative_recon_alg.py:214-230)."""
    g = copy.deepcopy(geo)
    g.sVoxel = np.asarray(g.sVoxel, dtype=np.float64).copy()
    g.sVoxel[1:] = g.sVoxel[1:] * 1.1
    g.sVoxel[0] = max(g.sDetector[0], g.sVoxel[0])
    g.nVoxel = np.array([2, 2, 2])
    g.dVoxel = g.sVoxel / g.nVoxel
    W = ax(np.ones((2, 2, 2), dtype=np.float32), prep(g, angles), "Siddon")
    W[W <= min(np.asarray(geo.dVoxel) / 2)] = np.inf
    return 1.0 / W


def weights_v(geo, angles, blocks):
    """V per block: mean over z of Atb(ones) over a shrunk volume (iterative_recon_alg.py:232-270);
    for a parallel beam, the number of views of the block (:265-267)."""
    if getattr(geo, "mode", "cone") == "parallel":
        ny, nx = int(geo.nVoxel[1]), int(geo.nVoxel[2])
        return np.stack([np.full((ny, nx), float(len(idx))) for idx in blocks])
    out = []
    for idx in blocks:
        g = sub_geometry(geo, idx)
        s = np.asarray(g.sVoxel, dtype=np.float64)
        g.sVoxel = s * np.max(s[1:] / np.linalg.norm(s[1:])) * 0.9
        g.dVoxel = g.sVoxel / np.asarray(g.nVoxel)
        ones = np.ones((len(idx), int(g.nDetector[0]), int(g.nDetector[1])), dtype=np.float32)
        out.append(atb(ones, prep(g, angles[idx]), "FDK").mean(axis=0))
    V = np.stack(out)
    V[V == 0.0] = np.inf
    return V


def os_sart(proj, geo, angles, niter, blocksize=20, lmbda=1.0, noneg=True, init=None):
    """iterative_recon_alg.py:326-401 with ordered subsets."""
    geo = copy.deepcopy(geo)
    geo.check_geo(angles)
    ang = geo.angles
    blocks = ordered_blocks(len(ang), blocksize)
    W = weights_w(geo, ang)
    V = weights_v(geo, ang, blocks)
    res = np.zeros(tuple(int(v) for v in geo.nVoxel)) if init is None else np.array(init, dtype=np.float64)
    proj = np.asarray(proj, dtype=np.float64)
    for _ in range(niter):
        for j, idx in enumerate(blocks):
            g = prep(sub_geometry(geo, idx), ang[idx])
            r = W[idx] * (proj[idx] - ax(res.astype(np.float32), g, "Siddon"))
            res = res + lmbda / V[j] * atb(r.astype(np.float32), g, "FDK")
            if noneg:
                res = np.maximum(res, 0)
    return res


# -- CGLS --------------------------------------------------------------------------
def cgls(proj, geo, angles, niter):
    """krylov_subspace_algorithms.py:45-97 (no restart branch), FP64 norms."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    x = np.zeros(tuple(int(v) for v in geo.nVoxel))
    r = proj - ax(x.astype(np.float32), g, "Siddon")
    p = atb(r.astype(np.float32), g, "matched")
    gamma = float(np.sum(p * p))
    res_norms = []
    for _ in range(niter):
        q = ax(p.astype(np.float32), g, "Siddon")
        alpha = gamma / float(np.sum(q * q))
        x = x + alpha * p
        res_norms.append(float(np.linalg.norm(proj - ax(x.astype(np.float32), g, "Siddon"))))
        r = r - alpha * q
        s = atb(r.astype(np.float32), g, "matched")
        gamma1 = float(np.sum(s * s))
        p = s + (gamma1 / gamma) * p
        gamma = gamma1
    return x, res_norms


# -- ASD-POCS ----------------------------------------------------------------------
def asd_pocs(proj, geo, angles, niter, alpha=0.002, alpha_red=0.95, rmax=0.95, tviter=20, lmbda=1.0,
             lmbda_red=1.0, epsilon=None, blocksize=1, trace=None, snapshots=None):
    """pocs_algorithms.py:105-169: a SART sweep over subsets of ``blocksize``
    views (1 for ASD-POCS) with the non-negativity clip, then ``tviter`` TV
    steps of size dtvg; dtvg = alpha * dp after the first sweep, reduced by
    alpha_red when dg > rmax * dp and dd > epsilon; beta = lmbda reduced by
    lmbda_red per iteration (it stops the loop below 0.005 but never scales the
    SART update, as in the reference).

    ``trace`` (a list) receives (n_iter, dd, dp, dg, c, dtvg) per iteration,
    dtvg being the step the TV call used; ``snapshots`` (a list) receives
    (pre-TV image, post-TV image) per iteration."""
    geo = copy.deepcopy(geo)
    geo.check_geo(angles)
    ang = geo.angles
    gfull = prep(geo, ang)
    proj = np.asarray(proj, dtype=np.float64)
    blocks = ordered_blocks(len(ang), blocksize)
    W = weights_w(geo, ang)
    V = weights_v(geo, ang, blocks)
    if epsilon is None:
        epsilon = 0.2 * np.linalg.norm(ax(fdk(proj, geo, angles)[0].astype(np.float32), gfull) - proj)
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    beta, n_iter = lmbda, 0
    dtvg = None
    while True:
        res_prev = res.copy()
        for j, idx in enumerate(blocks):
            g = prep(sub_geometry(geo, idx), ang[idx])
            r = W[idx] * (proj[idx] - ax(res.astype(np.float32), g, "Siddon"))
            res = np.maximum(res + lmbda / V[j] * atb(r.astype(np.float32), g, "FDK"), 0)
        n_iter += 1
        dd = np.linalg.norm(ax(res.astype(np.float32), gfull) - proj)
        dp_vec = res - res_prev
        dp = np.linalg.norm(dp_vec)
        if n_iter == 1:
            dtvg = alpha * dp
        res_prev = res.copy()
        used = dtvg
        res = pocs_tv(res, dtvg, tviter)
        if snapshots is not None:
            snapshots.append((res_prev, res.copy()))
        dg_vec = res - res_prev
        dg = np.linalg.norm(dg_vec)
        if dg > rmax * dp and dd > epsilon:
            dtvg *= alpha_red
        beta *= lmbda_red
        c = float(np.dot(dg_vec.ravel(), dp_vec.ravel())) / max(dg * dp, 1e-6)
        if trace is not None:
            trace.append((n_iter, float(dd), float(dp), float(dg), c, float(used)))
        if (c < -0.99 and dd <= epsilon) or beta < 0.005 or n_iter >= niter:
            return res


# -- the ROF denoiser's callers: SART-TV, OS-SART-TV, FISTA, ISTA -------------------
def im3ddenoise(img, tviter, lmbda):
    """utilities/im_3d_denoise.py:7-28: shift the image to minimum 0, scale it to
    maximum 1, ROF-denoise (tv_proximal, ``tviter`` iterations, ``lmbda``), undo
    the scaling and the shift."""
    img = np.asarray(img, dtype=np.float64)
    lo = img.min()
    x = img - lo
    hi = x.max()
    return tvdenoise(x / hi, lmbda, tviter) * hi + lo


def sart_tv(proj, geo, angles, niter, blocksize=1, tviter=50, tvlambda=50.0, lmbda=1.0, noneg=True):
    """art_family_algorithms.py:65-108 (SART_TV: blocksize 1) and :111-153
    (OSSART_TV: blocksize 20 by default): every iteration one (OS-)SART sweep
    of iterative_recon_alg.py:326-401, then im3ddenoise(res, tviter, tvlambda)."""
    geo = copy.deepcopy(geo)
    geo.check_geo(angles)
    ang = geo.angles
    blocks = ordered_blocks(len(ang), blocksize)
    W = weights_w(geo, ang)
    V = weights_v(geo, ang, blocks)
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    proj = np.asarray(proj, dtype=np.float64)
    for _ in range(niter):
        for j, idx in enumerate(blocks):
            g = prep(sub_geometry(geo, idx), ang[idx])
            r = W[idx] * (proj[idx] - ax(res.astype(np.float32), g, "Siddon"))
            res = res + lmbda / V[j] * atb(r.astype(np.float32), g, "FDK")
            if noneg:
                res = np.maximum(res, 0)
        res = im3ddenoise(res, tviter, tvlambda)
    return res


def fista(proj, geo, angles, niter, hyper=2.0e8, tviter=20, tvlambda=0.1, fista_p=1.0, fista_q=1.0, noneg=True,
          momentum=True):
    """ista_algorithms.py:113-182 (FISTA) and, with ``momentum`` False,
    :185-209 (ISTA).  One block of all views: res += (2 / hyper) Atb_matched(proj
    - Ax_interpolated(res)) (no W, V), clipped at 0; FISTA then denoises,
    x = im3ddenoise(res, tviter, 1 / (2 tvlambda / hyper)), and extrapolates,
    t' = (p + sqrt(q + 4 t^2)) / 2, res = x + (t - 1) / t' (x - x_old), with
    x_old starting at the initial image; ISTA sets res = im3ddenoise(res, 20,
    1 / (2 * 0.1 / hyper)) (its lambda is the fixed 0.1, its TV iterations 20)."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    bm = 1.0 / hyper
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    x_rec = res.copy()
    t = 1.0
    for _ in range(niter):
        res = res + bm * 2 * atb((proj - ax(res.astype(np.float32), g, "interpolated")).astype(np.float32), g,
                                 "matched")
        if noneg:
            res = np.maximum(res, 0)
        if not momentum:
            res = im3ddenoise(res, 20, 1.0 / (2 * bm * 0.1))
            continue
        x_old = x_rec
        x_rec = im3ddenoise(res, tviter, 1.0 / (2 * bm * tvlambda))
        t_old = t
        t = (fista_p + np.sqrt(fista_q + 4 * t ** 2)) / 2
        res = x_rec + (t_old - 1) / t * (x_rec - x_old)
    return res


# -- LSQR and MLEM -------------------------------------------------------------------
def lsqr(proj, geo, angles, niter, restart=True):
    """krylov_subspace_algorithms.py:101-188, Paige and Saunders' LSQR with
    Siddon Ax and matched Atb over all views.  After every iteration the data
    residual |proj - Ax(res)| is taken; when it rises the step is undone
    (res -= (phi / rho) w of the previous step) and the iteration ends (the
    reference re-initialises and then leaves its loop).  Returns (res,
    residual norms)."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    A = lambda x: ax(x.astype(np.float32), g, "Siddon")  # noqa: E731
    At = lambda y: atb(y.astype(np.float32), g, "matched")  # noqa: E731
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    u = proj - A(res)
    beta = np.linalg.norm(u)
    u = u / beta
    phibar = beta
    v = At(u)
    alpha = np.linalg.norm(v)
    v = v / alpha
    rhobar = alpha
    w = v.copy()
    norms = []
    for i in range(niter):
        u = A(v) - alpha * u
        beta = np.linalg.norm(u)
        u = u / beta
        v = At(u) - beta * v
        alpha = np.linalg.norm(v)
        v = v / alpha
        rho = np.sqrt(rhobar ** 2 + beta ** 2)
        c, s = rhobar / rho, beta / rho
        theta = s * alpha
        rhobar = -c * alpha
        phi = c * phibar
        phibar = s * phibar
        res = res + (phi / rho) * w
        w_prev = w
        w = v - (theta / rho) * w
        norms.append(float(np.linalg.norm(proj - A(res))))
        if i > 0 and norms[-1] > norms[-2]:
            res = res - (phi / rho) * w_prev
            break
    return res, norms


def mlem(proj, geo, angles, niter, init=None):
    """statistical_algorithms.py:11-55: res = 1 (or init); W = Atb_matched(1),
    infinite where <= 0; every iteration res *= Atb_matched(proj /
    Ax_interpolated(res)) / W (a zero forward projection divides by infinity),
    negatives clipped to 0 before and after."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(v) for v in geo.nVoxel)
    res = np.ones(shape) if init is None else np.array(init, dtype=np.float64)
    W = atb(np.ones(proj.shape, dtype=np.float32), g, "matched")
    W[W <= 0.0] = np.inf
    res[res < 0.0] = 0.0
    for _ in range(niter):
        den = ax(res.astype(np.float32), g, "interpolated")
        den[den == 0.0] = np.inf
        res = res * (atb((proj / den).astype(np.float32), g, "matched") / W)
        res[res < 0.0] = 0.0
    return res


def lsmr(proj, geo, angles, niter, lmbda=1.0):
    """krylov_subspace_algorithms.py:299-439, Fong and Saunders' LSMR with
    damping ``lmbda`` (the reference passes its lmbda, 1 by default), Siddon
    Ax and matched Atb over all views; the data residual is taken after every
    iteration and a rise undoes the step and ends the iteration, as in LSQR.
    Returns (res, residual norms)."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    A = lambda x: ax(x.astype(np.float32), g, "Siddon")  # noqa: E731
    At = lambda y: atb(y.astype(np.float32), g, "matched")  # noqa: E731
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    u = proj - A(res)
    beta = np.linalg.norm(u)
    u = u / beta
    v = At(u)
    alpha = np.linalg.norm(v)
    v = v / alpha
    alphabar, zetabar = alpha, alpha * beta
    rho, rhobar, cbar, sbar = 1.0, 1.0, 1.0, 0.0
    h, hbar = v.copy(), np.zeros_like(v)
    norms = []
    for i in range(niter):
        u = A(v) - alpha * u
        beta = np.linalg.norm(u)
        u = u / beta
        v = At(u) - beta * v
        alpha = np.linalg.norm(v)
        v = v / alpha
        alphahat = np.sqrt(alphabar ** 2 + lmbda ** 2)
        rhopre = rho
        rho = np.sqrt(alphahat ** 2 + beta ** 2)
        c, s = alphahat / rho, beta / rho
        theta = s * alpha
        alphabar = c * alpha
        thetabar = sbar * rho
        rhobarpre = rhobar
        rhobar = np.sqrt((cbar * rho) ** 2 + theta ** 2)
        cbar = cbar * rho / rhobar
        sbar = theta / rhobar
        zeta = cbar * zetabar
        zetabar = -sbar * zetabar
        hbar = h - (thetabar * rho / (rhopre * rhobarpre)) * hbar
        step = zeta / (rho * rhobar)
        res = res + step * hbar
        h = v - (theta / rho) * h
        norms.append(float(np.linalg.norm(proj - A(res))))
        if i > 0 and norms[-1] > norms[-2]:
            res = res - step * hbar
            break
    return res, norms


def gmres(proj, geo, angles, niter, variant="ba", backprojector="matched", restart=True):
    """krylov_subspace_algorithms.py:570-695: GMRES on the right- (AB) or
    left-preconditioned (BA) system, Siddon Ax and B = matched Atb (or FDK).
    BA: Arnoldi on B A over volumes from r0 = B(proj); AB: on A B over
    projections from r0 = proj, the image x = sum_i B(w_i) y_i.  After
    iteration k the data residual |proj - A x_k| is taken; when it rises at
    k = 1 (or at any k without ``restart``) that iterate is returned, as the
    reference does; after the last iteration x = sum_i w_i y_i (BA)."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    A = lambda x: ax(x.astype(np.float32), g, "Siddon")  # noqa: E731
    if backprojector == "FDK":
        B = lambda y: fdk(y, geo, angles)[0]  # noqa: E731
    else:
        B = lambda y: atb(y.astype(np.float32), g, "matched")  # noqa: E731
    r = B(proj) if variant == "ba" else proj.copy()
    beta = np.linalg.norm(r)
    w = [r / beta]
    z = []  # AB: B(w_i)
    H = np.zeros((niter + 1, niter))
    norms = []

    def image(y):
        basis = w if variant == "ba" else z
        return sum(yi * bi for yi, bi in zip(y, basis))

    y = None
    for k in range(niter):
        if variant == "ba":
            q = B(A(w[k]))
        else:
            z.append(B(w[k]))
            q = A(z[k])
        for i in range(k + 1):
            H[i, k] = float(np.sum(q * w[i]))
            q = q - H[i, k] * w[i]
        H[k + 1, k] = np.linalg.norm(q)
        w.append(q / H[k + 1, k])
        e1 = np.zeros(k + 2)
        e1[0] = beta
        y = np.linalg.lstsq(H[:k + 2, :k + 1], e1, rcond=None)[0]
        x = image(y)
        norms.append(float(np.linalg.norm(proj - A(x))))
        if k > 0 and norms[-1] > norms[-2] and (k == 1 or not restart):
            # the entry point returns the image the loop never assigned (the
            # zero start): krylov_subspace_algorithms.py:623-626, 688-691,
            # iterative_recon_alg.py:470-475
            return np.zeros_like(x), norms
    return image(y), norms


def hybrid_lsqr(proj, geo, angles, niter, lmbda=1.0, restart=True, elementwise_reorth=True):
    """krylov_subspace_algorithms.py:190-297: Golub-Kahan bidiagonalisation
    (Siddon Ax, matched Atb) whose projected problem min |B_k y - beta e1|^2 +
    lmbda^2 |y|^2 is solved through the SVD of B_k; x = V_k y.  The reference's
    "reorthogonalisation" against the earlier vectors j < i - 1 subtracts the
    elementwise product (V_j * v) * V_j (not a projection); it is kept as it
    computes (``elementwise_reorth`` False leaves it out, the textbook
    bidiagonalisation).  The data residual of each iterate is taken; a rise at
    the second iteration (or at any, without ``restart``) returns that
    iterate."""
    g = prep(geo, angles)
    proj = np.asarray(proj, dtype=np.float64)
    A = lambda x: ax(x.astype(np.float32), g, "Siddon")  # noqa: E731
    At = lambda y: atb(y.astype(np.float32), g, "matched")  # noqa: E731
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    u = proj - A(res)
    beta = np.linalg.norm(u)
    u = u / beta
    U, V = [u], []
    Bm = np.zeros((niter, niter + 1))
    rhs = np.zeros(niter + 1)
    rhs[0] = beta
    norms = []
    x = res
    for i in range(niter):
        v = At(u)
        if i > 0:
            v = v - beta * V[i - 1]
        for j in range(i - 1 if elementwise_reorth else 0):
            v = v - (V[j] * v) * V[j]
        alpha = np.linalg.norm(v)
        v = v / alpha
        V.append(v)
        u = A(v) - alpha * u
        for j in range(i - 1 if elementwise_reorth else 0):
            u = u - (U[j] * u) * U[j]
        beta = np.linalg.norm(u)
        u = u / beta
        U.append(u)
        Bm[i, i] = alpha
        Bm[i, i + 1] = beta
        Uk, Sk, Vk = np.linalg.svd(Bm[:i + 1, :i + 2].T)
        yhat = Sk * (Uk.T @ rhs[:i + 2])[:i + 1] / (Sk ** 2 + lmbda ** 2)
        y = Vk.T @ yhat
        x = res + sum(yj * vj for yj, vj in zip(y, V))
        norms.append(float(np.linalg.norm(proj - A(x))))
        if i > 0 and norms[-1] > norms[-2] and (i == 1 or not restart):
            # returned from the loop, not assigned to the image that the entry
            # point returns (the zero start): :285-289, iterative_recon_alg.py:470-475
            return np.zeros_like(x), norms
    return x, norms


# -- PCSD ----------------------------------------------------------------------------
def pcsd(proj, geo, angles, niter, blocksize=1, tviter=20, lmbda=1.0, lmbda_red=1.0, epsilon=None, delta=None,
         trace=None):
    """pocs_algorithms.py:434-646 (PCSD, AwPCSD with ``delta``, and their OS
    forms with ``blocksize``): every iteration, when d_p^2 > epsilon (d_p the
    data residual of the current image; the reference compares the squared
    norm with epsilon) an (OS-)SART sweep; then the TV step (pocs_tv, or
    aw_pocs_tv with ``delta``) of size dtvg = 1 in the first iteration and
    d_p / d_p1 after it, d_p1 the data residual of the first pre-TV image;
    beta = lmbda reduced by lmbda_red; the loop stops when c < -0.99 and the
    post-sweep residual is within epsilon, when beta < 0.005, or after
    iteration niter + 1 (the reference tests n_iter > niter).  ``trace``
    receives (n_iter, d_p, dd, dp, dtvg, dg, c) per iteration."""
    geo = copy.deepcopy(geo)
    geo.check_geo(angles)
    ang = geo.angles
    gfull = prep(geo, ang)
    proj = np.asarray(proj, dtype=np.float64)
    blocks = ordered_blocks(len(ang), blocksize)
    W = weights_w(geo, ang)
    V = weights_v(geo, ang, blocks)
    A = lambda x: ax(x.astype(np.float32), gfull)  # noqa: E731
    if epsilon is None:
        epsilon = 0.2 * np.linalg.norm(A(fdk(proj, geo, angles)[0]) - proj)
    res = np.zeros(tuple(int(v) for v in geo.nVoxel))
    beta, n_iter, d_p1 = lmbda, 0, 1.0
    while True:
        res_prev = res.copy()
        n_iter += 1
        d_p = np.linalg.norm(A(res) - proj)
        if d_p ** 2 > epsilon:
            for j, idx in enumerate(blocks):
                g = prep(sub_geometry(geo, idx), ang[idx])
                r = W[idx] * (proj[idx] - ax(res.astype(np.float32), g, "Siddon"))
                res = np.maximum(res + lmbda / V[j] * atb(r.astype(np.float32), g, "FDK"), 0)
        dd = np.linalg.norm(A(res) - proj)
        dp_vec = res - res_prev
        dp = np.linalg.norm(dp_vec)
        dtvg = 1.0 if n_iter == 1 else d_p / d_p1
        res_prev = res.copy()
        res = pocs_tv(res, dtvg, tviter) if delta is None else aw_pocs_tv(res, dtvg, tviter, delta)
        dg_vec = res - res_prev
        dg = np.linalg.norm(dg_vec)
        if n_iter == 1:
            d_p1 = np.linalg.norm(A(res_prev) - proj)
        beta *= lmbda_red
        c = float(np.dot(dg_vec.ravel(), dp_vec.ravel())) / max(dg * dp, 1e-6)
        if trace is not None:
            trace.append((n_iter, float(d_p), float(dd), float(dp), float(dtvg), float(dg), c))
        if (c < -0.99 and dd <= epsilon) or beta < 0.005 or n_iter > niter:
            return res


# -- IRN-TV-CGLS ---------------------------------------------------------------------
def irn_diffs(x):
    """The differences of krylov_subspace_algorithms.py:456-476, one array per
    axis: d_i = x_i - x_{i+1} for i < n - 2, while the last two entries along
    that axis keep x itself (the reference overwrites [0:-2] of a copy)."""
    out = []
    for a in range(3):
        d = np.array(x, dtype=np.float64, copy=True)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, -2), slice(1, -1)
        d[tuple(lo)] = d[tuple(lo)] - np.asarray(x, np.float64)[tuple(hi)]
        out.append(d)
    return out


def irn_weights(x):
    """(|D x|^2 + 1e-6)^(-1/4) (krylov_subspace_algorithms.py:456-465)."""
    d = irn_diffs(x)
    return (d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + 1e-6) ** -0.25


def irn_lx(W, x):
    """Lx (:467-476): W D x, three arrays."""
    return [W * d for d in irn_diffs(x)]


def irn_ltx(W, y):
    """Ltx (:478-496), as the reference computes it: per axis w = W y_a, then
    out_0 = w_0, out_i = w_i - w_{i-1} for 0 < i < n - 2, out_{n-2} = w_{n-2},
    out_{n-1} = -w_{n-2} (not the transpose of irn_lx on its last two rows)."""
    total = 0.0
    for a in range(3):
        w = W * np.asarray(y[a], np.float64)
        o = w.copy()
        mid, prev, last, second = ([slice(None)] * 3 for _ in range(4))
        mid[a], prev[a], last[a], second[a] = slice(1, -2), slice(0, -3), -1, -2
        o[tuple(mid)] = w[tuple(mid)] - w[tuple(prev)]
        o[tuple(last)] = -w[tuple(second)]
        total = total + o
    return total


def irn_cycles(b, x0, A, At, weights, lx, ltx, niter, niter_outer, lmbda):
    """The loop of krylov_subspace_algorithms.py:498-565 over given operators:
    each outer cycle builds W from the current image, then runs niter CGLS
    steps on [A; sqrt(lmbda) W D] x = [b; 0] from res0, the image before the
    previous cycle's last update (the reference's res0 is reassigned inside
    the inner loop).  FP64 throughout; returns (x, per-step ||b - A x||)."""
    sl = float(np.sqrt(lmbda))
    res = np.asarray(x0, dtype=np.float64)
    res0 = res
    norms = []
    for _ in range(niter_outer):
        W = weights(res)
        res = res0
        r1 = b - A(res)
        r2 = [-sl * t for t in lx(W, res)]
        p = At(r1) + sl * ltx(W, r2)
        gamma = float(np.sum(p * p))
        for _ in range(niter):
            res0 = res
            q1 = A(p)
            q2 = [sl * t for t in lx(W, p)]
            alpha = gamma / (float(np.sum(q1 * q1)) + sum(float(np.sum(t * t)) for t in q2))
            res = res + alpha * p
            norms.append(float(np.linalg.norm(b - A(res))))
            r1 = r1 - alpha * q1
            r2 = [u - alpha * v for u, v in zip(r2, q2)]
            s = At(r1) + sl * ltx(W, r2)
            gamma1 = float(np.sum(s * s))
            p = s + (gamma1 / gamma) * p
            gamma = gamma1
    return res, norms


def irn_tv_cgls(proj, geo, angles, niter, lmbda=1.0, niter_outer=4, init=None):
    """IRN-TV-CGLS (krylov_subspace_algorithms.py:441-567) with Siddon forward
    and matched backprojection."""
    g = prep(geo, angles)
    b = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(v) for v in geo.nVoxel)
    x0 = np.zeros(shape) if init is None else np.asarray(init, np.float64).reshape(shape)
    return irn_cycles(b, x0, lambda x: ax(np.asarray(x, np.float32), g, "Siddon"),
                      lambda y: atb(np.asarray(y, np.float32), g, "matched"),
                      irn_weights, irn_lx, irn_ltx, niter, niter_outer, lmbda)


# -- hybrid fLSQR-TV -----------------------------------------------------------------
def hybrid_flsqr_tv(proj, geo, angles, niter, lmbda=1.0, restart=True, init=None, trace=None):
    """Hybrid flexible LSQR with TV (krylov_subspace_algorithms.py:698-860) in
    FP64.  The constant image that best fits the data is split off (xA0);
    each step backprojects u, orthogonalises it against the earlier v,
    preconditions it by two scipy lsqr solves (50 iterations at most, scipy's
    default tolerances, as the reference calls them) with the operators
    irn_ltx / irn_lx, projects it, extends the U basis with A z, and solves
    min ||M_k^T y - beta e_1||^2 + lmbda^2 ||R y||^2, R the triangular factor
    of the stacked W D z_j.  Returns (x, norms, start): x the algorithm's
    image; start the image the reference's entry point returns, which its
    loop never assigns (iterative_recon_alg.py:470-475)."""
    from scipy.linalg import qr
    from scipy.sparse.linalg import LinearOperator, lsqr as sp_lsqr

    g = prep(geo, angles)
    b = np.asarray(proj, dtype=np.float64)
    shape = tuple(int(v) for v in geo.nVoxel)
    N = int(np.prod(shape))
    res = np.zeros(shape) if init is None else np.asarray(init, np.float64).reshape(shape)
    A = lambda x: ax(np.asarray(x, np.float32), g, "Siddon")
    At = lambda y: atb(np.asarray(y, np.float32), g, "matched")
    u = b - A(res)
    k = A(np.ones(shape) / np.sqrt(N))
    c = 1.0 / (np.sqrt(N) * float(np.sum(k * k)))
    xA0 = np.full(shape, c * float(np.sum(k * u)))
    kv = c * At(k)
    u = u - A(xA0)
    normr = float(np.linalg.norm(u))
    if trace is not None:
        trace.update(r0=u.copy(), xA0=xA0, Z=[])
    u = u / normr
    W = np.ones(shape) if res.max() == 0 else irn_weights(res)
    lx = lambda x: np.concatenate([t.ravel() for t in irn_lx(W, np.reshape(x, shape))])
    ltx = lambda y: irn_ltx(W, list(np.reshape(y, (3,) + shape))).ravel()
    L1 = LinearOperator((N, 3 * N), matvec=ltx, rmatvec=lx, dtype=np.float64)
    L2 = LinearOperator((3 * N, N), matvec=lx, rmatvec=ltx, dtype=np.float64)
    U, V, Z = [u], [], []
    M = np.zeros((niter, niter + 1))
    rhs = np.zeros(niter + 1)
    rhs[0] = normr
    norms = []
    x = res
    for i in range(niter):
        v = At(u)
        for vj in V:
            v = v - float(np.sum(vj * v)) * vj
        v = v / np.linalg.norm(v)
        z = v - kv * float(np.sum(v))
        aux = sp_lsqr(L1, z.ravel(), iter_lim=50)[0]
        z = sp_lsqr(L2, aux, iter_lim=50)[0].reshape(shape)
        z = z - float(np.sum(kv * z))
        V.append(v)
        Z.append(z)
        u = A(z)
        for j, uj in enumerate(U):
            M[i, j] = float(np.sum(uj * u))
            u = u - M[i, j] * uj
        M[i, i + 1] = np.linalg.norm(u)
        u = u / M[i, i + 1]
        U.append(u)
        R = qr(np.stack([lx(zj) for zj in Z], 1), mode="economic")[1]
        y = np.linalg.lstsq(np.vstack([M[:i + 1, :i + 2].T, lmbda * R]),
                            np.concatenate([rhs[:i + 2], np.zeros(i + 1)]), rcond=None)[0]
        x = res + sum(yj * zj for yj, zj in zip(y, Z)) + xA0
        norms.append(float(np.linalg.norm(b - A(x))))
        if i > 0 and norms[-1] > norms[-2] and (i == 1 or not restart):
            break
    if trace is not None:
        trace.update(Z=Z, W=W)
    return x, norms, res
