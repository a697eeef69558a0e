"""The reference's mTPQ step restated in numpy on top of the CPU oracle's H·v
(test helper): MakeIniVec normalisation (src/MakeIniVec.c:95-116), FirstMultiply
(src/FirstMultiply.c:41-85), Multiply (src/Multiply.c:42-66), expec_energy_flct
(src/expec_energy_flct.c:126-157,429-500) and inv_temp (src/CalcByTPQ.c:235)."""

import numpy as np


def tpq_restated(hv, v_init, steps, lv, Ns, sz_of=None, num_of=None):
    """Rows (inv_temp, energy, <H^2>, norm, 2<Sz>, 4<Sz^2>) of one sample; with
    ``num_of`` (the electron number of each state) a seventh column <N>."""
    rows = []
    first_norm = np.linalg.norm(v_init)
    v1 = v_init / first_norm
    h = hv(v1)

    def obs(v):
        w = np.abs(v) ** 2
        o = (np.sum(w * sz_of), np.sum(w * sz_of ** 2)) if sz_of is not None else (0.0, 0.0)
        return o + ((np.sum(w * num_of),) if num_of is not None else ())

    o = obs(v1)
    rows.append((0.0, np.vdot(v1, h).real, np.vdot(h, h).real, first_norm, *o))
    for s in range(1, steps):
        v0 = lv * v1 - h / Ns
        gnorm = np.linalg.norm(v0)
        v1 = v0 / gnorm
        h = hv(v1)
        E = np.vdot(v1, h).real
        o = obs(v1)
        rows.append((2.0 * s / Ns / (lv - E / Ns), E, np.vdot(h, h).real, gnorm, *o))
    return np.array(rows)


def ctpq_restated(hv, v_init, steps, dtau, coef, obs=None):
    """Rows (inv_temp, energy, <H^2>, norm, *obs(v)) of one canonical TPQ sample:
    MakeIniVec normalisation, then per step MultiplyForCanonicalTPQ
    (src/Multiply.c:152-205: v <- sum_{m<=coef} (-dtau/2)^m/m! H^m v, reusing
    the measured H v, normalised; its norm is global_norm) and
    expec_energy_flct; inv_temp grows by dtau per step
    (src/CalcByCanonicalTPQ.c:140-250)."""
    obs = obs or (lambda v: ())
    first_norm = np.linalg.norm(v_init)
    v1 = v_init / first_norm
    h = hv(v1)
    rows = [(0.0, np.vdot(v1, h).real, np.vdot(h, h).real, first_norm, *obs(v1))]
    inv_temp = 0.0
    for _ in range(1, steps):
        t = -0.5 * dtau
        acc = v1 + t * h
        cur = h
        for m in range(2, coef + 1):
            t *= (-0.5 * dtau) / m
            cur = hv(cur)
            acc = acc + t * cur
        gnorm = np.linalg.norm(acc)
        v1 = acc / gnorm
        h = hv(v1)
        inv_temp += dtau
        rows.append((inv_temp, np.vdot(v1, h).real, np.vdot(h, h).real, gnorm, *obs(v1)))
    return np.array(rows)
