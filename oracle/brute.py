"""50-digit brute force of the Plummer direct sum -- TEST INFRASTRUCTURE ONLY.

Pins the C oracle independently of floating point: every FP64 input is converted
to ``decimal.Decimal`` exactly, displacements and squares are exact at 60 digits,
the square root is Decimal's correctly rounded ``sqrt``.  Pure Python loops, so
only for tiny cases (<= a few hundred pairs).  Same kernel and sign convention as
``oracle_eval_plummer`` (d = x_s - x_t, eps2 = eps_s^2 + eps_t^2, zero pairs skipped).
"""
from decimal import Decimal, getcontext


def eval_plummer(sx, sy, sz, sm, se, tx, ty, tz, te, G, digits=60):
    getcontext().prec = digits
    D = Decimal
    Gd = D(float(G))
    out = []
    for i in range(len(tx)):
        xi, yi, zi = D(float(tx[i])), D(float(ty[i])), D(float(tz[i]))
        e2t = D(float(te[i])) ** 2 if te is not None else D(0)
        gx = gy = gz = gp = D(0)
        for j in range(len(sx)):
            dx = D(float(sx[j])) - xi
            dy = D(float(sy[j])) - yi
            dz = D(float(sz[j])) - zi
            e2 = (D(float(se[j])) ** 2 if se is not None else D(0)) + e2t
            r2 = dx * dx + dy * dy + dz * dz + e2
            if r2 == 0:
                continue
            rinv = 1 / r2.sqrt()
            mr = D(float(sm[j])) * rinv
            w = mr * rinv * rinv
            gx += w * dx
            gy += w * dy
            gz += w * dz
            gp += mr
        out.append((Gd * gx, Gd * gy, Gd * gz, -Gd * gp))
    return out
