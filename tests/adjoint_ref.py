"""FP64 torch restatement of the density and momentum-energy loops over explicit pair
lists, for differentiating with torch.autograd.  Its forward values are checked against
the oracle (tests/test_adjoint_ref.py), its gradients against central differences of
the oracle, and it then serves as the reference for the hand-written GPU adjoints.

Follows hydro_ve/xmass_kern.hpp:50-79 (density), hydro_std/momentum_energy_kern.hpp:
12-134 (momentum-energy), sph/table_lookup.hpp:13-26 (table-linear W) and the legacy
periodic fold sfc/box.hpp:217-238."""
import numpy as np
import torch

NUM_INTERVALS = 19999


def pairs_from_lists(nb, nc, ngmax):
    """(i, j) int64 arrays of the stored pairs: the first min(nc - 1, ngmax) entries of each row."""
    nb = np.asarray(nb).reshape(len(nc), ngmax)
    cnt = np.clip(np.asarray(nc, np.int64) - 1, 0, ngmax)
    mask = np.arange(ngmax)[None, :] < cnt[:, None]
    i = np.repeat(np.arange(len(nc), dtype=np.int64), cnt)
    j = nb[mask].astype(np.int64)
    return torch.from_numpy(i), torch.from_numpy(j)


def table_w(wh, v):
    dx = 2.0 / NUM_INTERVALS
    inv = 1.0 / dx
    idx = torch.floor(v.detach() * inv).long()
    inside = idx < NUM_INTERVALS
    idc = idx.clamp(max=NUM_INTERVALS - 1)
    w0 = wh[idc]
    dw = (wh[idc + 1] - w0) * inv
    return torch.where(inside, w0 + dw * (v - idc.double() * dx), torch.zeros_like(v))


def fold(r, thr, length, periodic):
    if not periodic:
        return r
    return torch.where(r > thr, r - length, torch.where(r < -thr, r + length, r))


def box_info(box):
    lens = (box.xmax - box.xmin, box.ymax - box.ymin, box.zmax - box.zmin)
    per = (box.bx == 1, box.by == 1, box.bz == 1)
    return lens, per


def offsets(x, y, z, i, j, thr, box):
    lens, per = box_info(box)
    rx = fold(x[i] - x[j], thr, lens[0], per[0])
    ry = fold(y[i] - y[j], thr, lens[1], per[1])
    rz = fold(z[i] - z[j], thr, lens[2], per[2])
    return rx, ry, rz


def density(x, y, z, h, m, i, j, wh, K, box):
    rx, ry, rz = offsets(x, y, z, i, j, 2.0 * h[i], box)
    dist = torch.sqrt(rx * rx + ry * ry + rz * rz)
    rho0 = m + torch.zeros_like(m).index_add(0, i, m[j] * table_w(wh, dist / h[i]))
    return K / h ** 3 * rho0


def momentum_energy(x, y, z, f, i, j, wh, K, box):
    """f: dict of vx, vy, vz, h, m, rho, p, c, c11..c33 (double tensors).  Returns ax, ay, az, du."""
    h, m, rho, p, c = f["h"], f["m"], f["rho"], f["p"], f["c"]
    rx, ry, rz = offsets(x, y, z, i, j, 2.0 * h[i], box)
    dist = torch.sqrt(rx * rx + ry * ry + rz * rz)
    vx, vy, vz = (f[k][i] - f[k][j] for k in ("vx", "vy", "vz"))
    hiInv, hjInv = 1.0 / h[i], 1.0 / h[j]
    Wi = hiInv ** 3 * table_w(wh, dist * hiInv)
    Wj = hjInv ** 3 * table_w(wh, dist * hjInv)

    def sym(q, a, b, cc):
        return (f["c11"][q] * a + f["c12"][q] * b + f["c13"][q] * cc,
                f["c12"][q] * a + f["c22"][q] * b + f["c23"][q] * cc,
                f["c13"][q] * a + f["c23"][q] * b + f["c33"][q] * cc)

    Ai = sym(i, rx, ry, rz)
    Aj = sym(j, rx, ry, rz)
    w = (rx * vx + ry * vy + rz * vz) / dist
    ci, cj = c[i], c[j]
    visc = torch.where(w < 0, -((1.0 + 1.0) * 0.25 * (ci + cj) - 2.0 * w) * w, torch.zeros_like(w))
    Pi = 0.5 * visc
    mj_pro_i = m[j] * p[i] / (rho[i] * rho[i])
    mi_roi = m[i] / rho[i]
    mj_roj_Wj = m[j] / rho[j] * Wj
    a = Wi * (mj_pro_i + Pi * mi_roi)
    b = mj_roj_Wj * (p[j] / rho[j] + Pi)
    a2 = Wi * (2.0 * mj_pro_i + Pi * mi_roi)
    b2 = Pi * mj_roj_Wj
    zero = torch.zeros_like(h)
    mom = [zero.index_add(0, i, a * Ai[k] + b * Aj[k]) for k in range(3)]
    en = zero.index_add(0, i, vx * (a2 * Ai[0] + b2 * Aj[0]) + vy * (a2 * Ai[1] + b2 * Aj[1]) +
                        vz * (a2 * Ai[2] + b2 * Aj[2]))
    return K * mom[0], K * mom[1], K * mom[2], -K * 0.5 * en


def tt(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64))
