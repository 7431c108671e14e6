"""Independent pins of the oracle's swarm init and particle update.

`or_init` restates InitQPSO (main.cu:72-198) and `or_update_particles`
restates UpdateParticlesEquationKernel (main.cu:858-1060, with the quaternion
helpers of main.h:208-413).  The restatement here is written from those lines
in a different form from the oracle's C:

* every random draw is addressed explicitly as T[p + s*N], with s listed per
  statement of the reference (init: s = 0..11; update of iteration `it`:
  s = 12 + 12*it + k, k = 0..11 in statement order);
* the fp32 multiply-add sites are evaluated exactly (rationals) and rounded
  once, the inertia product as the reference's double product rounded to
  float;
* the rotation part is float64 quaternion algebra in scalar/vector form:
  q~ = q_target (x) q^-1, the log map through atan2(|v|, w), and the update
  as right multiplication by the axis-angle quaternion of angle |w|/2.

Translations must match bit for bit; the rotation part within the error
that the fp32 evaluation of acos(q~0) and the products allows (bounded below).
`test_mutants_are_caught` rebuilds the oracle with each of several plausible
mistakes (conjugate order, draw order, fp32 inertia, c1/c2 swapped, ...) and
checks that these tests reject every one.
"""
import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(REPO, "oracle", "qpso_oracle.c")
CAM = (525.0, 525.0, 319.5, 239.5)

# statement order of InitQPSO (main.cu:98-145): draw index s of each value
INIT_DRAW = {"pos_tx": 0, "vel_tx": 1, "pos_ty": 2, "vel_ty": 3, "pos_tz": 4, "vel_tz": 5,
             "wpert_q1": 6, "wpert_q2": 7, "wpert_q3": 8, "vel_q1": 9, "vel_q2": 10, "vel_q3": 11}


# ------------------------------------------------------------- exact fp32
def f32_round(x: Fraction) -> np.float32:
    """Correctly rounded (nearest, ties to even) binary32 value of a rational."""
    f = np.float32(float(x))
    best, bd = None, None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        d = abs(Fraction(float(c)) - x)
        if bd is None or d < bd or (d == bd and (int(c.view(np.uint32)) & 1) == 0):
            best, bd = c, d
    return np.float32(best)


def fma32(a, b, c):
    return f32_round(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


# ------------------------------------------------------------- float64 quaternions
def hamilton(a, b):
    """a (x) b for scalar-first quaternions, float64, rows of (..., 4)."""
    aw, av, bw, bv = a[..., 0], a[..., 1:], b[..., 0], b[..., 1:]
    w = aw * bw - np.sum(av * bv, axis=-1)
    v = aw[..., None] * bv + bw[..., None] * av + np.cross(av, bv)
    return np.concatenate([w[..., None], v], axis=-1)


def conj(q):
    return q * np.array([1.0, -1.0, -1.0, -1.0])


def top_half(q):
    return np.where(q[..., :1] < 0, -q, q)


def log_map(qt):
    """Rotation vector of a top-half quaternion: 2 atan2(|v|, w) v/|v| (0 when |v| <= 1e-7)."""
    v = qt[..., 1:]
    n = np.linalg.norm(v, axis=-1)
    ang = 2.0 * np.arctan2(n, qt[..., 0])
    scale = np.where(n > 1e-7, ang / np.where(n > 0, n, 1.0), 0.0)
    return v * scale[..., None]


def axis_angle(axis_vec, angle):
    """Unit quaternion of `angle` about the direction of axis_vec."""
    n = np.linalg.norm(axis_vec, axis=-1)
    u = axis_vec / np.where(n > 0, n, 1.0)[..., None]
    return np.concatenate([np.cos(angle / 2)[..., None], np.sin(angle / 2)[..., None] * u], axis=-1)


def kinematic_step(q, omega):
    """QuatKinematicsPSOStateVect (main.h:374-413) as q (x) exp of angle |omega|/2, normalised, q0 >= 0."""
    out = hamilton(q, axis_angle(omega, np.linalg.norm(omega, axis=-1) / 2.0))
    out = out / np.linalg.norm(out, axis=-1)[..., None]
    return top_half(out)


# ------------------------------------------------------------- restatements
def init_expected(T, N, lo, hi):
    """InitQPSO (main.cu:98-145) for every particle: exact translations, float64 quaternion."""
    def r(p, name):
        return T[p + INIT_DRAW[name] * N]

    pos = np.zeros((7, N), np.float32)
    vel = np.zeros((7, N), np.float32)
    q = np.zeros((N, 4))
    for p in range(N):
        for d, ax in enumerate("xyz"):
            pos[d, p] = fma32(np.float32(hi[d]) - np.float32(lo[d]), r(p, "pos_t" + ax), lo[d])
            vel[d, p] = fma32(np.float32(6.0), r(p, "vel_t" + ax), np.float32(-3.0))  # -vdot_ + 2 vdot_ r
        wp = np.array([fma32(np.float32(20.0), r(p, f"wpert_q{k}"), np.float32(-10.0)) for k in (1, 2, 3)])
        for k in (1, 2, 3):
            vel[3 + k, p] = fma32(np.float32(20.0), r(p, f"vel_q{k}"), np.float32(-10.0))
        qinit = np.array([np.float32(0.7071), np.float32(0.7071), 0.0, 0.0], np.float64)
        q[p] = kinematic_step(qinit, wp.astype(np.float64))
    return pos, vel, q


def update_expected(T, N, s0, w, c1, c2, lo, hi, pos, vel, pbest, gbest):
    """UpdateParticlesEquationKernel (main.cu:921-1031) for every particle.

    Returns the exact new translations and translational velocities, and the
    float64 rotational velocities."""
    P = pos.copy()
    V = vel.copy()
    for p in range(N):
        for d in range(3):
            r1 = np.float32(np.float32(c1) * T[p + (s0 + 2 * d) * N])
            r2 = np.float32(np.float32(c2) * T[p + (s0 + 2 * d + 1) * N])
            iv = np.float32(np.float64(w) * np.float64(vel[d, p]))  # psoW_ is a double
            a = fma32(r1, np.float32(pbest[d, p] - pos[d, p]), iv)
            v = fma32(r2, np.float32(gbest[d] - pos[d, p]), a)
            x = np.float32(pos[d, p] + v)
            if x < lo[d]:
                x, v = np.float32(lo[d]), np.float32(0.0)
            elif x > hi[d]:
                x, v = np.float32(hi[d]), np.float32(0.0)
            P[d, p], V[d, p] = x, v
    q = pos[3:].T.astype(np.float64)
    wc = log_map(top_half(hamilton(pbest[3:].T.astype(np.float64), conj(q))))
    ws = log_map(top_half(hamilton(np.broadcast_to(gbest[3:7].astype(np.float64), q.shape), conj(q))))
    omega = np.zeros((N, 3))
    for i in range(3):
        idx = np.arange(N)
        r1 = (np.float32(c1) * T[idx + (s0 + 6 + 2 * i) * N]).astype(np.float32)
        r2 = (np.float32(c2) * T[idx + (s0 + 7 + 2 * i) * N]).astype(np.float32)
        iw = (np.float64(w) * vel[4 + i].astype(np.float64)).astype(np.float32)
        omega[:, i] = iw.astype(np.float64) + r1 * wc[:, i] + r2 * ws[:, i]
    return P, V, omega


# ------------------------------------------------------------- checks (return failures)
def _params(N, w, c1, c2, lo, hi, iters=5):
    return O.make_params(N, iters, w, c1, c2, lo, hi, [100, 100, 200, 200], 1000, CAM)


def _random_table(N, rows, seed):
    return np.random.default_rng(seed).random(N * rows, dtype=np.float64).astype(np.float32)


def check_init(L, N=1024):
    lo = [np.float32(-0.31), np.float32(-0.27), np.float32(0.52)]
    hi = [np.float32(0.29), np.float32(0.33), np.float32(1.13)]
    par = _params(N, 0.7, 1.3, 1.7, lo, hi)
    T = _random_table(N, 512, 11)
    pos = np.zeros((7, N), np.float32); vel = np.zeros_like(pos); pb = np.zeros_like(pos)
    pbf = np.zeros(N, np.float32); gb = np.zeros(8, np.float32)
    L.or_init(ctypes.byref(par), T.ctypes.data, pos.ctypes.data, vel.ctypes.data, pb.ctypes.data,
              pbf.ctypes.data, gb.ctypes.data)
    epos, evel, eq = init_expected(T, N, lo, hi)
    fails = []
    if not np.array_equal(pos[:3].view(np.uint32), epos[:3].view(np.uint32)):
        fails.append("init translation")
    if not (np.array_equal(vel[:3].view(np.uint32), evel[:3].view(np.uint32)) and np.all(vel[3] == 0)
            and np.array_equal(vel[4:].view(np.uint32), evel[4:].view(np.uint32))):
        fails.append("init velocity")
    if np.max(np.abs(pos[3:].T.astype(np.float64) - eq)) > 2e-6:
        fails.append("init quaternion")
    if not (np.array_equal(pos, pb) and np.all(pbf == np.float32(1e10)) and gb[7] == np.float32(1e10)
            and np.all(gb[:7] == 0)):
        fails.append("init personal/global best")
    return fails


def make_update_state(N, seed):
    """States where pbest != gbest != pos with q~0 of both targets in [0.17, 0.99]."""
    rng = np.random.default_rng(seed)
    lo = np.array([-0.3, -0.3, 0.5], np.float32)
    hi = np.array([0.3, 0.3, 1.1], np.float32)

    def rand_unit(n):
        q = rng.normal(size=(n, 4))
        return q / np.linalg.norm(q, axis=1)[:, None]

    def rand_angle(n):
        return rng.uniform(0.3, 2.8, n)

    gq = rand_unit(1)[0]
    pq = hamilton(np.broadcast_to(gq, (N, 4)), axis_angle(rng.normal(size=(N, 3)), rand_angle(N)))
    pq = np.where(rng.random(N)[:, None] < 0.5, pq, -pq)
    pq = top_half(pq)
    bq = top_half(hamilton(pq, axis_angle(rng.normal(size=(N, 3)), rand_angle(N))))
    pos = np.zeros((7, N), np.float32)
    pos[:3] = rng.uniform(lo[:, None], hi[:, None], (3, N))
    pos[3:] = pq.T
    pbest = np.zeros_like(pos)
    pbest[:3] = np.clip(pos[:3] + rng.normal(0, 0.05, (3, N)), lo[:, None], hi[:, None])
    pbest[3:] = bq.T
    vel = np.zeros_like(pos)
    vel[:3] = rng.uniform(-0.1, 0.1, (3, N))
    big = rng.random(N) < 0.1  # some leave the bounds: clamp and zero velocity
    vel[:3, big] = rng.uniform(-3, 3, (3, int(big.sum())))
    vel[4:] = rng.uniform(-2, 2, (3, N))
    gbest = np.zeros(8, np.float32)
    gbest[:3] = rng.uniform(lo, hi)
    gbest[3:7] = gq
    gbest[7] = 0.25
    return lo, hi, pos, vel, pbest, gbest


def check_update(L, N=1024, it=3, w=0.7, c1=1.3, c2=1.7):
    lo, hi, pos, vel, pbest, gbest = make_update_state(N, 5 + it)
    par = _params(N, w, c1, c2, lo, hi, iters=it + 1)
    T = _random_table(N, 12 * (it + 2), 40 + it)
    s0 = 12 + 12 * it
    P, V = pos.copy(), vel.copy()
    L.or_update_particles(ctypes.byref(par), T.ctypes.data, s0, P.ctypes.data, V.ctypes.data,
                          pbest.ctypes.data, gbest.ctypes.data)
    eP, eV, omega = update_expected(T, N, s0, w, c1, c2, lo, hi, pos, vel, pbest, gbest)
    fails = []
    if not np.array_equal(P[:3].view(np.uint32), eP[:3].view(np.uint32)):
        fails.append("update translation")
    if not np.array_equal(V[:3].view(np.uint32), eV[:3].view(np.uint32)):
        fails.append("update translational velocity")
    if not np.all(V[3] == 0):
        fails.append("update vel q0")
    # rotation: the oracle evaluates acos(q~0) in fp32 from fp32 products; with q~0 <= 0.99
    # (|acos'| <= 7) that is a few 1e-7 rad per term, times rho <= 1.9: measured <= 1.5e-6
    if np.max(np.abs(V[4:].T.astype(np.float64) - omega)) > 1e-5:
        fails.append("update angular velocity")
    q_new = kinematic_step(pos[3:].T.astype(np.float64), V[4:].T.astype(np.float64))
    if np.max(np.abs(P[3:].T.astype(np.float64) - q_new)) > 1e-6:  # measured <= 1.4e-7
        fails.append("update quaternion")
    return fails


def _tiny_scene():
    """Two faces of a square 60 mm wide at 0.8 m facing the camera, and its observed depth."""
    h = np.float32(0.03)
    A = [(-h, -h, 0.8), (h, -h, 0.8), (h, h, 0.8)]
    B = [(-h, -h, 0.8), (h, h, 0.8), (-h, h, 0.8)]
    faces = [A, B]
    F = 2
    soa = np.zeros(9 * F, np.float32)
    for f, tri in enumerate(faces):
        for j, (x, y, z) in enumerate(tri):
            soa[f + j * F], soa[3 * F + f + j * F], soa[6 * F + f + j * F] = x, y, z
    obs = np.zeros((480, 640), np.uint16)
    obs[220:260, 300:340] = 800
    return soa, obs


def check_solve_draws(L, N=8, iters=3):
    """or_solve reads exactly the draws s < 12 (I + 1): init 0..11, iteration it 12 + 12 it .. 23 + 12 it."""
    soa, obs = _tiny_scene()
    par = O.make_params(N, iters, 0.7, 1.3, 1.7, [-0.05, -0.05, 0.75], [0.05, 0.05, 0.85],
                        [300, 220, 340, 260], 1600, CAM)
    T = _random_table(N, 12 * (iters + 1) + 24, 77)
    used = 12 * (iters + 1)

    def run(table):
        res = {"fit": np.zeros((iters, N), np.float32), "vel": np.zeros((7, N), np.float32),
               "pos": np.zeros((7, N), np.float32)}
        o = O.SolveOut()
        o.fit_trace, o.vel, o.pos = res["fit"].ctypes.data, res["vel"].ctypes.data, res["pos"].ctypes.data
        rc = L.or_solve(soa.ctypes.data, 2, obs.ctypes.data, ctypes.byref(par), table.ctypes.data, 1,
                        ctypes.byref(o))
        assert rc == 0
        return res

    def same(a, b):
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))

    base = run(T)
    fails = []
    beyond = T.copy()
    beyond[used * N:] = np.nan
    r = run(beyond)
    if not (same(r["fit"], base["fit"]) and same(r["vel"], base["vel"]) and same(r["pos"], base["pos"])):
        fails.append("solve reads draws beyond 12 (I + 1)")
    last = T.copy()
    last[0 + (used - 1) * N] = np.nan  # particle 0's last draw: rho2 of the last rotation axis
    r = run(last)
    if not (same(r["fit"], base["fit"]) and np.isnan(r["vel"][6, 0]) and same(r["pos"][:3], base["pos"][:3])):
        fails.append("solve: last draw not the final update's rotation rho2")
    first = T.copy()
    first[0 + 12 * N] = np.nan  # particle 0's first update draw: rho1 of tx after iteration 0
    r = run(first)
    if not (same(r["fit"][0], base["fit"][0]) and np.isnan(r["pos"][0, 0])):
        fails.append("solve: draw 12 not the first update's tx rho1")
    return fails


# ------------------------------------------------------------- tests
def test_init_draws_and_values():
    assert check_init(O.lib()) == []


@pytest.mark.parametrize("it,w,c1,c2", [(0, 0.7, 1.3, 1.7), (3, 0.3, 1.0, 1.0), (7, 0.5, 1.9, 0.6)])
def test_update_matches_restatement(it, w, c1, c2):
    assert check_update(O.lib(), it=it, w=w, c1=c1, c2=c2) == []


def test_solve_consumes_draws_in_order():
    assert check_solve_draws(O.lib()) == []


def test_restatement_agrees_on_quaternion_algebra():
    # the restatement's own pieces against closed forms: q (x) q^-1 = identity,
    # log(exp) = identity, a right multiplication composes angles about one axis
    rng = np.random.default_rng(3)
    q = rng.normal(size=(50, 4)); q /= np.linalg.norm(q, axis=1)[:, None]
    assert np.allclose(hamilton(q, conj(q)), [1, 0, 0, 0], atol=1e-14)
    w = rng.uniform(-2, 2, (50, 3))
    assert np.allclose(log_map(top_half(axis_angle(w, np.linalg.norm(w, axis=1)))), w, atol=1e-12)
    ax = rng.normal(size=(50, 3))
    a1, a2 = rng.uniform(0, 1, 50), rng.uniform(0, 1, 50)
    assert np.allclose(hamilton(axis_angle(ax, a1), axis_angle(ax, a2)), axis_angle(ax, a1 + a2), atol=1e-14)
    # kinematic_step(q, omega) rotates by |omega| / 2 (mTs_ = 0.5)
    e = kinematic_step(np.tile([1.0, 0, 0, 0], (50, 1)), w)
    assert np.allclose(2 * np.arccos(np.clip(e[:, 0], -1, 1)), np.linalg.norm(w, axis=1) / 2, atol=1e-7)


# Each mutant replaces text of oracle/qpso_oracle.c (every `old` must occur
# exactly `count` times) with a plausible mistake.
MUTANTS = {
    "conjugate order p (x) pb*": [("qprod_conj(targets[k], &P[3], qt);", "qprod_conj(&P[3], targets[k], qt);", 1)],
    "rho2 drawn before rho1": [(
        "const float r1 = pr->c1 * RT(p, s); ++s;\n            const float r2 = pr->c2 * RT(p, s); ++s;\n"
        "            const float iV",
        "const float r2 = pr->c2 * RT(p, s); ++s;\n            const float r1 = pr->c1 * RT(p, s); ++s;\n"
        "            const float iV", 1)],
    "rotation draws before translation": [
        ("int s = rand_idx;", "int s = rand_idx + 6;", 1),
        ("V[3] = 0.0f;\n        for (int i = 0; i < 3; ++i) {",
         "V[3] = 0.0f;\n        s = rand_idx;\n        for (int i = 0; i < 3; ++i) {", 1)],
    "fp32 inertia": [("(float)(pr->w * (double)V[d])", "((float)pr->w * V[d])", 1),
                     ("(float)(pr->w * (double)V[4 + i])", "((float)pr->w * V[4 + i])", 1)],
    "c1 and c2 swapped": [("pr->c1 * RT", "pr->cX * RT", 2), ("pr->c2 * RT", "pr->c1 * RT", 2),
                          ("pr->cX * RT", "pr->c2 * RT", 2)],
    "social term toward pbest": [("{&B[3], &gbest[3]}", "{&B[3], &B[3]}", 1)],
    "no top-half cast of q~": [("qprod_conj(targets[k], &P[3], qt);\n            top_half(qt);",
                               "qprod_conj(targets[k], &P[3], qt);", 1)],
    "init pos x3 then vel x3": [(
        "            P[d] = fmaf(pr->x_hi[d] - pr->x_lo[d], RT(p, s), pr->x_lo[d]); ++s;\n"
        "            V[d] = fmaf(6.0f, RT(p, s), -3.0f); ++s; /* -vdot_ + 2 vdot_ r */\n        }",
        "            P[d] = fmaf(pr->x_hi[d] - pr->x_lo[d], RT(p, s), pr->x_lo[d]); ++s;\n        }\n"
        "        for (int d = 0; d < 3; ++d) {\n"
        "            V[d] = fmaf(6.0f, RT(p, s), -3.0f); ++s;\n        }", 1)],
    "init angular velocity drawn before the perturbation": [
        ("for (int d = 0; d < 3; ++d) { wp[d] = fmaf(20.0f, RT(p, s), -10.0f); ++s; }",
         "for (int d = 0; d < 3; ++d) { wp[d] = fmaf(20.0f, RT(p, s + 3), -10.0f); }", 1),
        ("for (int d = 0; d < 3; ++d) { V[4 + d] = fmaf(20.0f, RT(p, s), -10.0f); ++s; }",
         "for (int d = 0; d < 3; ++d) { V[4 + d] = fmaf(20.0f, RT(p, s), -10.0f); ++s; } s += 3;", 1)],
    "update draws start at 12 it": [("12 + 12 * it, pos, vel", "12 * it, pos, vel", 1)],
    "translation clamp keeps the velocity": [("P[d] = pr->x_lo[d]; V[d] = 0.0f;", "P[d] = pr->x_lo[d];", 1)],
}


def _build_mutant(tmp, name, edits):
    src = open(ORACLE_SRC).read()
    for old, new, count in edits:
        assert src.count(old) == count, f"mutant {name!r}: {old!r} occurs {src.count(old)} times"
        src = src.replace(old, new)
    tag = "".join(c if c.isalnum() else "_" for c in name)
    cpath = os.path.join(tmp, f"mut_{tag}.c")
    so = os.path.join(tmp, f"libmut_{tag}.so")
    with open(cpath, "w") as f:
        f.write(src)
    subprocess.check_call(["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fwrapv",
                           "-fopenmp", "-shared", "-I", os.path.join(REPO, "oracle"), "-o", so, cpath, "-lm"])
    return O.load(so)


def test_mutants_are_caught(tmp_path):
    missed = []
    for name, edits in MUTANTS.items():
        L = _build_mutant(str(tmp_path), name, edits)
        fails = check_init(L, N=256) + check_update(L, N=256) + check_solve_draws(L)
        if not fails:
            missed.append(name)
    assert missed == [], f"mutants not caught: {missed}"
