"""Host FP64 recomputation of the solver's per-depth true residual from the
exported operator inputs (test helper; one neighbour slot at a time, so the
2M-point depth-9 case needs no [M][27] temporaries).

The operator entry of row o and neighbour slot q is, as the reference
assembles it (LaplacianSolver.cu:40-70),
    L = dot_x dot_y dot_z (d2dot_x + d2dot_y + d2dot_z)
over the 1-D inner-product tables indexed by the two nodes' function indices,
stored as float when |L| > 1e-6 and dropped otherwise."""
import numpy as np


def decode(enc, D):
    s1 = 1 << (D + 1)
    return enc & (s1 - 1), (enc >> (D + 1)) & (s1 - 1), enc >> (2 * (D + 1))


def true_relres(rec, D):
    """||b - A x|| / ||b|| per depth in FP64, b the exported divergence and x
    the exported solution, A rebuilt from the exported tables."""
    res = (1 << (D + 1)) - 1
    dot = rec.export("TABLE_DOT").reshape(res, res)
    d2 = rec.export("TABLE_D2DOT").reshape(res, res)
    enc = rec.export("NODE_ENCIDX").astype(np.int64).reshape(-1)
    neigh = rec.export("NODE_NEIGHS").reshape(-1, 27)
    b = rec.export("DIVERGENCE").astype(np.float64)
    x = rec.export("DX").astype(np.float64)
    base = rec.export("LEVEL_BASE").astype(np.int64)
    fx, fy, fz = decode(enc, D)
    ax = np.zeros_like(b)
    for q in range(27):
        nb = neigh[:, q]
        ok = nb >= 0
        o = np.nonzero(ok)[0]
        j = nb[ok].astype(np.int64)
        sx, sy, sz = (fx[o], fx[j]), (fy[o], fy[j]), (fz[o], fz[j])
        e = dot[sx] * dot[sy] * dot[sz] * (d2[sx] + d2[sy] + d2[sz])
        v = np.where(np.abs(e) > float(np.float32(1e-6)), e.astype(np.float32).astype(np.float64), 0.0)
        ax[o] += v * x[j]
    out = []
    for d in range(D + 1):
        a, c = base[d], base[d + 1]
        nb_ = np.linalg.norm(b[a:c])
        out.append(float(np.linalg.norm(b[a:c] - ax[a:c]) / nb_) if nb_ > 0 else 0.0)
    return out
