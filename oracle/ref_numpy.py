"""ORACLE — test infrastructure only.

A numpy restatement of ``reference_fa2`` of the reference's test suite
(tests/testthat/helper-reference.R:3-100): a vectorised FP64 transcription of
Gephi's ForceAtlas2 iteration with exact repulsion and every option. The
reference tests use it as the independent check of the C code
(tests/testthat/test-algorithm.R:1-61); here it pins the C oracle the same
way. Written from the R text, operation by operation.
"""
import numpy as np


def reference_fa2(n, edges, pos, iterations, scaling, gravity=1.0,
                  strong=False, linlog=False, outbound=False, adjust=False,
                  size=1.0, w=None, jt_user=1.0, fixed=None):
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    a, b = e[:, 0], e[:, 1]
    w = np.ones(len(a)) if w is None else np.asarray(w, np.float64)
    fixed = np.zeros(n, bool) if fixed is None else np.asarray(fixed, bool)
    size = np.broadcast_to(np.asarray(size, np.float64), (n,))
    mass = 1.0 + np.bincount(np.concatenate([a, b]), minlength=n)  # :14
    x = np.array(pos[:, 0], np.float64)
    y = np.array(pos[:, 1], np.float64)
    dx = np.zeros(n)
    dy = np.zeros(n)
    speed, eff = 1.0, 1.0
    mm = np.outer(mass, mass)
    ss = np.add.outer(size, size)
    for _ in range(iterations):
        odx, ody = dx, dy
        # repulsion (:27-40)
        xd = np.subtract.outer(x, x)
        yd = np.subtract.outer(y, y)
        d = np.sqrt(xd ** 2 + yd ** 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            if adjust:
                d = d - ss
                f = np.where(d > 0, scaling * mm / d ** 2,
                             np.where(d < 0, 100 * scaling * mm, 0.0))
            else:
                f = np.where(d > 0, scaling * mm / d ** 2, 0.0)
        np.fill_diagonal(f, 0.0)
        dx = (xd * f).sum(1)
        dy = (yd * f).sum(1)
        # gravity (:43-46)
        d0 = np.sqrt(x ** 2 + y ** 2)
        if strong:
            g = gravity * mass
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                g = np.where(d0 > 0, gravity * mass / d0, 0.0)
        dx = dx - x * g
        dy = dy - y * g
        # attraction, edge by edge (:49-67)
        coef = mass.mean() if outbound else 1.0
        for k in range(len(a)):
            i, j = a[k], b[k]
            ex = x[i] - x[j]
            ey = y[i] - y[j]
            fe = -coef * w[k]
            if adjust or linlog:
                de = np.sqrt(ex ** 2 + ey ** 2)
                if adjust:
                    de = de - size[i] - size[j]
                if de <= 0:
                    continue
                if linlog:
                    fe = fe * np.log(1 + de) / de
            if outbound:
                fe = fe / mass[i]
            dx[i] += ex * fe
            dy[i] += ey * fe
            dx[j] -= ex * fe
            dy[j] -= ey * fe
        # speed (:70-85)
        swing_i = mass * np.sqrt((odx - dx) ** 2 + (ody - dy) ** 2)
        swing = swing_i[~fixed].sum()
        tract = (0.5 * mass * np.sqrt((odx + dx) ** 2 + (ody + dy) ** 2))[~fixed].sum()
        est = 0.05 * np.sqrt(n)
        jt = jt_user * max(np.sqrt(est), min(10.0, est * tract / n ** 2))
        if swing / tract > 2:
            if eff > 0.05:
                eff *= 0.5
            jt = max(jt, jt_user)
        target = jt * eff * tract / swing
        if swing > jt * tract:
            if eff > 0.05:
                eff *= 0.7
        elif speed < 1000:
            eff *= 1.3
        speed = speed + min(target - speed, 0.5 * speed)
        # displacement (:88-97)
        if adjust:
            fac = 0.1 * speed / (1 + np.sqrt(speed * swing_i))
            df = np.sqrt(dx ** 2 + dy ** 2)
            with np.errstate(divide="ignore", invalid="ignore"):
                fac = np.where(df == 0, 0.0, np.minimum(fac * df, 10) / df)
        else:
            fac = speed / (1 + np.sqrt(speed * swing_i))
        fac = np.where(fixed, 0.0, fac)
        x = x + dx * fac
        y = y + dy * fac
    return np.stack([x, y], 1)


def update_speed(n, swing, tract, jt_user, speed, eff):
    """The speed block of reference_fa2 (:73-85) as a scalar function."""
    if not swing > 0:
        return speed, eff
    est = 0.05 * np.sqrt(n)
    jt = jt_user * max(np.sqrt(est), min(10.0, est * tract / n ** 2))
    if swing / tract > 2:
        if eff > 0.05:
            eff *= 0.5
        jt = max(jt, jt_user)
    target = jt * eff * tract / swing
    if swing > jt * tract:
        if eff > 0.05:
            eff *= 0.7
    elif speed < 1000:
        eff *= 1.3
    return speed + min(target - speed, 0.5 * speed), eff
