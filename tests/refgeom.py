"""World-frame cone-beam geometry for the oracle pins (test helper).

An independent statement of TIGRE's scanner model, in millimetres and world
coordinates (no voxel-index frames):
  - gantry rotation R = Rz(alpha) Ry(theta) Rz(psi); source at R (DSO, 0, 0);
  - detector pixel (u, v) (v = array row) sits at detector-plane coordinates
    y = dU (u + 1/2 - nU/2) + offU, z = dV (v + 1/2 - nV/2) + offV, the
    roll/pitch/yaw rotation Rz(roll) Ry(pitch) Rx(yaw) acting about the
    detector centre before the offsets; the plane is at x = -(DSD - DSO);
  - a centre-of-rotation correction moves source and detector by
    (-COR sin alpha, COR cos alpha, 0);
  - the volume is the box centred on offOrigin with size sVoxel.
Arrays follow the Python front end: volumes (z, y, x), detector (v, u).
"""
import numpy as np


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])


def rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


def view_params(g, i):
    """Per-view scalars of a prepared geometry (after check_geo/cast_to_single)."""
    a, t, p = [float(x) for x in np.asarray(g.angles)[i]]
    off = np.asarray(g.offOrigin, dtype=np.float64).reshape(-1, 3)[i]   # (z, y, x)
    offd = np.asarray(g.offDetector, dtype=np.float64).reshape(-1, 2)[i]  # (v, u)
    rot = np.asarray(g.rotDetector, dtype=np.float64).reshape(-1, 3)[i]   # (yaw, pitch, roll)
    DSO = float(np.asarray(g.DSO).reshape(-1)[i])
    DSD = float(np.asarray(g.DSD).reshape(-1)[i])
    COR = float(np.asarray(g.COR).reshape(-1)[i])
    return a, t, p, off[::-1].copy(), offd[::-1].copy(), rot[::-1].copy(), DSO, DSD, COR


def source_and_pixel(g, i, u, v):
    """Source and pixel-centre positions (x, y, z) in mm for view i, pixel (u, v)."""
    a, t, p, off_xyz, offd_uv, rot_rpy, DSO, DSD, COR = view_params(g, i)
    R = rz(a) @ ry(t) @ rz(p)
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    dV, dU = float(g.dDetector[0]), float(g.dDetector[1])
    yu = dU * (u + 0.5 - nU / 2.0)
    zv = dV * (v + 0.5 - nV / 2.0)
    D = rz(rot_rpy[0]) @ ry(rot_rpy[1]) @ rx(rot_rpy[2]) @ np.array([0.0, yu, zv])
    D = D + np.array([-(DSD - DSO), offd_uv[0], offd_uv[1]])
    shift = np.array([-COR * np.sin(a), COR * np.cos(a), 0.0])
    S = R @ np.array([DSO, 0.0, 0.0]) + shift
    P = R @ D + shift
    return S, P


def voxel_size(g):
    """dVoxel (x, y, z) in mm, as given (not sVoxel / nVoxel)."""
    return np.asarray(g.dVoxel, dtype=np.float64)[::-1]


def box(g, i):
    """Volume box (lo, hi) in mm, (x, y, z): the corner sits at offOrigin - sVoxel/2 and
    voxel k spans [lo + k dVoxel, lo + (k + 1) dVoxel]."""
    _, _, _, off_xyz, _, _, _, _, _ = view_params(g, i)
    s = np.asarray(g.sVoxel, dtype=np.float64)[::-1]
    n = np.asarray([int(x) for x in np.asarray(g.nVoxel)][::-1], dtype=np.float64)
    lo = off_xyz - s / 2
    return lo, lo + n * voxel_size(g)


def segment_box_length(S, P, lo, hi):
    """Length of the segment S->P inside the axis-aligned box [lo, hi]."""
    d = P - S
    t0, t1 = 0.0, 1.0
    for k in range(3):
        if d[k] == 0:
            if S[k] < lo[k] or S[k] > hi[k]:
                return 0.0
            continue
        a, b = (lo[k] - S[k]) / d[k], (hi[k] - S[k]) / d[k]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0) * float(np.linalg.norm(d))


def voxel_path_lengths(g, i, u, v):
    """Exact intersection length of the ray S->P with every voxel, (z, y, x) array."""
    S, P = source_and_pixel(g, i, u, v)
    lo, hi = box(g, i)
    n = [int(x) for x in np.asarray(g.nVoxel)][::-1]  # x, y, z
    dvox = voxel_size(g)
    out = np.zeros(n[::-1])
    for k in range(n[2]):
        for j in range(n[1]):
            for ii in range(n[0]):
                vlo = lo + dvox * np.array([ii, j, k])
                out[k, j, ii] = segment_box_length(S, P, vlo, vlo + dvox)
    return out


def project_point(g, i, X):
    """Continuous detector coordinates (u, v) (pixel-index units) of world point X (mm)."""
    a, t, p, off_xyz, offd_uv, rot_rpy, DSO, DSD, COR = view_params(g, i)
    R = rz(a) @ ry(t) @ rz(p)
    shift = np.array([-COR * np.sin(a), COR * np.cos(a), 0.0])
    Xl = R.T @ (X - shift)            # gantry frame
    S = np.array([DSO, 0.0, 0.0])
    M = rz(rot_rpy[0]) @ ry(rot_rpy[1]) @ rx(rot_rpy[2])
    c = np.array([-(DSD - DSO), offd_uv[0], offd_uv[1]])  # detector centre
    # solve S + s (Xl - S) = c + M (0, y, z)  for (s, y, z)
    A = np.stack([Xl - S, -M[:, 1], -M[:, 2]], axis=1)
    s, y, z = np.linalg.solve(A, c - S)
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    dV, dU = float(g.dDetector[0]), float(g.dDetector[1])
    return y / dU + nU / 2.0 - 0.5, z / dV + nV / 2.0 - 0.5


def voxel_centre(g, i, iz, iy, ix):
    lo, _ = box(g, i)
    return lo + voxel_size(g) * (np.array([ix, iy, iz]) + 0.5)
