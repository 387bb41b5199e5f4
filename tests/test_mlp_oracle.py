"""Pins the CPU oracle of the learned far-field correction: the backward pass
against finite differences, bf16 rounding against hand-worked values, the
targets against the closed-form quadrupole field, the second moments against
the parallel-axis theorem."""
import numpy as np

import mlp_oracle as MO


def _params(rng, scale=0.3):
    return rng.normal(size=MO.N_PARAMS) * scale


def test_layout():
    assert MO.N_PARAMS == 10448
    rng = np.random.default_rng(0)
    p = _params(rng)
    np.testing.assert_array_equal(MO.pack(MO.unpack(p)), p)


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to even (1.0);
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6... ties to even
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9,
                     -2.5, 3.0e-5], dtype=np.float32)
    r = MO.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0
    assert r[2] == 1.0 + 2.0 ** -6
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-5) <= 2.0 ** -8 * 3.0e-5


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x = rng.normal(size=(37, MO.IN))
    t = rng.normal(size=(37, MO.USED))
    g, L = MO.gradient(p, x, t, emulate=False)
    assert abs(L - MO.loss(p, x, t)) < 1e-12
    idx = rng.choice(MO.N_PARAMS, 60, replace=False)
    # include every layer's bias block
    idx = np.concatenate([idx, [1024, 1088 + 4096, 9408 + 1024 + 3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (MO.loss(pp, x, t) - MO.loss(pm, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_rounded_gradient_close_to_exact():
    rng = np.random.default_rng(2)
    p = _params(rng)
    x = rng.normal(size=(128, MO.IN))
    t = rng.normal(size=(128, MO.USED))
    g1, _ = MO.gradient(p, x, t, emulate=False)
    g2, _ = MO.gradient(p, x, t, emulate=True)
    assert np.linalg.norm(g1 - g2) < 0.05 * np.linalg.norm(g1)


def test_targets_reduce_to_quadrupole_far_away():
    # two equal bodies (no octupole): the residual is the quadrupole field to
    # O((s/d)^2); the gradient matches a finite difference of the field
    s = 0.01
    bx = np.array([s, -s]); by = np.array([0.3 * s, -0.3 * s]); bz = np.array([0.0, 0.0])
    bm = np.array([0.5, 0.5])
    com = np.zeros(3)
    Q = MO.second_moments(bx, by, bz, bm)
    g = np.array([0.7, -0.4, 0.5])
    hl = 2 * s
    x, t = MO.pair_inputs_targets(bx, by, bz, bm, com, 1.0, hl, Q, g, 0.0)
    r = com - g
    d = np.sqrt(r @ r)
    sres = hl * hl / d ** 4
    quad = MO.quadrupole_field(Q, r)
    res, J_back = MO.correction_from_outputs(t, com, 1.0, hl, g)
    np.testing.assert_allclose(res, quad, rtol=2e-3, atol=1e-3 * np.abs(quad).max())
    # gradient: central differences of the residual field
    eps = 1e-6
    src = np.stack([bx, by, bz], axis=1)

    def res(point):
        a, _ = MO.field_and_gradient(src, bm, point, 0.0)
        am, _ = MO.field_and_gradient(com[None, :], [1.0], point, 0.0)
        return a - am

    J = np.zeros((3, 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = eps
        J[:, j] = (res(g + e) - res(g - e)) / (2 * eps)
    np.testing.assert_allclose(J_back, J, rtol=1e-5, atol=1e-6 * np.abs(J).max())
    # inputs: size ratio and the moments in the cell frame
    assert abs(x[0] - hl / d) < 1e-15
    F = MO.frame(r / d)
    qp = F @ MO.sym_matrix(Q / hl ** 2) @ F.T
    np.testing.assert_allclose(x[1:7], [qp[0, 0], qp[1, 1], qp[2, 2], qp[0, 1], qp[0, 2],
                                        qp[1, 2]], atol=1e-14)


def test_frame_is_orthogonal_and_aligns():
    rng = np.random.default_rng(4)
    for u in list(rng.normal(size=(50, 3))) + [np.array([0, 0, 1.0]), np.array([0, 0, -1.0]),
                                               np.array([1.0, 0, 0]), np.array([0, 1e-9, -1])]:
        u = u / np.linalg.norm(u)
        F = MO.frame(u)
        np.testing.assert_allclose(F @ F.T, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(F @ u, [0, 0, 1], atol=1e-12)


def test_second_moments_parallel_axis():
    rng = np.random.default_rng(3)
    b = rng.normal(size=(40, 3))
    m = rng.uniform(0.1, 1.0, 40)
    Q = MO.second_moments(b[:, 0], b[:, 1], b[:, 2], m)
    tot = np.zeros(6)
    M = m.sum()
    c = (m[:, None] * b).sum(0) / M
    for part in (slice(0, 13), slice(13, 40)):
        bp, mp = b[part], m[part]
        Qp = MO.second_moments(bp[:, 0], bp[:, 1], bp[:, 2], mp)
        cp = (mp[:, None] * bp).sum(0) / mp.sum()
        dcv = cp - c
        tot += Qp + mp.sum() * np.array([dcv[0] ** 2, dcv[1] ** 2, dcv[2] ** 2,
                                          dcv[0] * dcv[1], dcv[0] * dcv[2], dcv[1] * dcv[2]])
    np.testing.assert_allclose(tot, Q, rtol=1e-12)
