"""Second, independent implementation of one DUGKS step (tests only).

Written from the method itself rather than from the oracle's structure:
whole-array numpy, and the face value of population q is the trilinear
interpolation of f-bar+ at the characteristic foot point x_face - c_q h,
expressed as a displacement from the cell centre in cell units (normal axis:
s/2 - h c_a/dx_a toward the face neighbour; tangential axes: -h c_b/dx_b
toward the upwind neighbour).  Uniform axes, trilinear faces only.  Method:
Guo et al., PRE 88 033305 (2013); stage structure of the reference
src/dugks.cu:322-391.
"""
import numpy as np

from conftest import C, OPP, RT, W


def coefficients(dt, nu):
    tau = nu / RT
    h = dt / 2
    return dict(tau=tau, h=h, bw0=3 * tau * h / (2 * tau + dt), bw1=(2 * tau - h) / (2 * tau + dt),
                bw2=3 * h / (2 * tau + dt), fw0=tau * h / (2 * tau + h),
                fw1=2 * tau / (2 * tau + h), fw2=h / (2 * tau + h))


def _feq(rho, u):
    # rho (...), u (3, ...) -> (19, ...)
    cu = np.tensordot(C, u, axes=(1, 0)) / RT
    uu = (u * u).sum(0) / RT
    return W.reshape((19,) + (1,) * rho.ndim) * rho * (1 + cu + 0.5 * (cu * cu - uu))


def _collide(g, F, h, w0, w1, w2):
    """g (19, ...): returns w1 g + (w2 + w0 (c - u).F/RT) feq(rho, u), u = j/rho + h F."""
    rho = g.sum(0)
    u = np.tensordot(C.T, g, axes=(1, 0)) / rho + h * F
    T = (np.tensordot(C, F, axes=(1, 0)) - (u * F).sum(0)) / RT
    return w1 * g + (w2 + w0 * T) * _feq(rho, u)


def step(f, periodic, dx, dt, nu, F=(0.0, 0.0, 0.0), Fcell=None):
    """f (nz, ny, nx, 19) -> f after one step.  Fcell: per-cell centre force
    (3, nz, ny, nx) replacing the ambient F at centres."""
    co = coefficients(dt, nu)
    h = co["h"]
    g = np.ascontiguousarray(np.transpose(f, (3, 2, 1, 0)))  # (19, nx, ny, nz)
    n = g.shape[1:]
    Fa = np.asarray(F, dtype=float).reshape(3, 1, 1, 1)
    Fc = Fa * np.ones((1,) + n) if Fcell is None else np.transpose(np.asarray(Fcell), (0, 3, 2, 1))
    fb = _collide(g, Fc, h, co["bw0"], co["bw1"], co["bw2"])
    E = np.zeros((19,) + tuple(m + 2 for m in n))
    E[:, 1:-1, 1:-1, 1:-1] = fb
    for a in range(3):
        def sl(m):
            idx = [slice(None)] * 4
            idx[a + 1] = m
            return tuple(idx)
        if periodic[a]:
            E[sl(0)] = E[sl(n[a])]
            E[sl(n[a] + 1)] = E[sl(1)]
        else:
            E[sl(0)] = 2 * E[sl(1)] - E[sl(2)]
            E[sl(n[a] + 1)] = 2 * E[sl(n[a])] - E[sl(n[a] - 1)]

    def at(q, off):
        return E[q, 1 + off[0]:1 + off[0] + n[0], 1 + off[1]:1 + off[1] + n[1],
                 1 + off[2]:1 + off[2] + n[2]]

    acc = 4.0 / 3.0 * fb - g / 3.0
    for a in range(3):
        for s in (-1, 1):
            ft = np.empty_like(g)
            for q in range(19):
                disp = [-h * C[q, b] / dx[b] for b in range(3)]
                disp[a] += 0.5 * s
                sgn = [int(np.sign(d)) for d in disp]
                val = 0.0
                for corner in range(8):
                    bits = [(corner >> b) & 1 for b in range(3)]
                    wgt = 1.0
                    for b in range(3):
                        wgt *= abs(disp[b]) if bits[b] else 1 - abs(disp[b])
                    if wgt == 0.0:
                        continue
                    val = val + wgt * at(q, [bits[b] * sgn[b] for b in range(3)])
                ft[q] = val
            if not periodic[a]:
                idx = [slice(None)] * 3
                idx[a] = 0 if s < 0 else n[a] - 1
                idx = tuple(idx)
                for q in range(19):
                    if C[q, a] * s < 0:
                        ft[(q,) + idx] = ft[(OPP[q],) + idx]
            ft = _collide(ft, Fa, h / 2, co["fw0"], co["fw1"], co["fw2"])
            acc -= s * (dt / dx[a]) * C[:, a].reshape(19, 1, 1, 1) * ft
    return np.transpose(acc, (3, 2, 1, 0))
