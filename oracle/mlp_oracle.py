"""CPU restatement of the learned far-field correction (test infrastructure:
the oracle for the MLP kernels of libcmhip.so; never used by the product).

No reference counterpart: the reference's Barnes-Hut is monopole-only
(src/CelestialMechanics.Physics/Solvers/BarnesHutBackend.cs:392-459); this
restates the definitions of the build's network and training targets:

* network: 16 -> 64 -> 64 -> 64 -> 16, ReLU on hidden layers, parameters
  per layer W[out][in] row-major then b[out] (layers 64x16, 64x64, 64x64,
  16x64, 10448 floats in total);
* rounding points of the GPU kernels (``emulate=True``): inputs, weights,
  hidden activations and backward deltas dZ rounded to bf16 (round to nearest
  even), everything else exact here (float64) where the GPU accumulates in
  fp32;
* loss: mean over samples and the 8 used outputs of (y - t)^2;
* targets of one (group centre g, cell) pair: exact softened field of the
  cell's bodies at g and its gradient, minus the monopole's, in units of
  M h^2 / d^4 (field) and M h^2 / d^5 (gradient), d = |com - g|.

Pinned by: finite differences of ``loss`` against ``gradient`` (float64, no
rounding); the targets against the closed-form quadrupole field for a
two-body cell far away (``quadrupole_field``).
"""
import numpy as np

IN, HID, OUT, USED = 16, 64, 16, 8
SHAPES = [(HID, IN), (HID, HID), (HID, HID), (OUT, HID)]
N_PARAMS = sum(o * i + o for o, i in SHAPES)   # 10448


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, k = [], 0
    for o, i in SHAPES:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def forward(params, x, emulate=True):
    """Outputs [n][16] and the activations (inputs of each layer)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params)
    a = r(np.asarray(x, dtype=np.float32).reshape(-1, IN))
    acts = [a]
    for li, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        if li < 3:
            a = r(np.maximum(z, 0.0))
            acts.append(a)
        else:
            return z, acts
    raise AssertionError


def loss(params, x, t, emulate=False):
    y, _ = forward(params, x, emulate)
    d = y[:, :USED] - np.asarray(t, dtype=np.float64).reshape(-1, USED)
    return float(np.mean(d * d))


def gradient(params, x, t, emulate=True):
    """d loss / d params (the GPU's order and rounding points) and the loss."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params)
    y, acts = forward(params, x, emulate)
    n = len(y)
    t = np.asarray(t, dtype=np.float64).reshape(-1, USED)
    d = y[:, :USED] - t
    dz = np.zeros_like(y)
    dz[:, :USED] = d * (2.0 / (USED * n))
    dz = r(dz)
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        a_in = acts[li]
        gW = dz.T @ a_in
        gb = dz.sum(axis=0)
        grads[li] = (gW, gb)
        if li > 0:
            da = dz @ r(W)
            dz = r(np.where(a_in > 0.0, da, 0.0))
    return pack(grads), float(np.mean(d * d))


def field_and_gradient(src, m, g, eps2):
    """Softened field sum m r / (r^2 + eps2)^1.5 at g (r = src - g) and its
    gradient d a_i / d x_j (float64)."""
    r = np.asarray(src, dtype=np.float64) - np.asarray(g, dtype=np.float64)
    D = np.einsum("ij,ij->i", r, r) + eps2
    i3 = D ** -1.5
    i5 = D ** -2.5
    m = np.asarray(m, dtype=np.float64)
    a = (m * i3) @ r
    G = 3.0 * np.einsum("k,ki,kj->ij", m * i5, r, r) - np.eye(3) * np.sum(m * i3)
    return a, G


def frame(u):
    """Orthogonal F with F u = +z: diag(1, 1, -s) (I - w w^T / (1 + |u_z|)),
    w = u + s z, s = sign(u_z) (s = 1 for u_z = 0)."""
    u = np.asarray(u, dtype=np.float64)
    s = 1.0 if u[2] >= 0.0 else -1.0
    w = u + np.array([0.0, 0.0, s])
    H = np.eye(3) - np.outer(w, w) / (1.0 + abs(u[2]))
    return np.diag([1.0, 1.0, -s]) @ H


def sym_matrix(v):
    """(xx, yy, zz, xy, xz, yz) -> 3x3."""
    return np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])


def pair_inputs_targets(bx, by, bz, bm, com, M, hl, Q, g, eps2):
    """Inputs (16) and targets (8) of one (group centre g, cell) pair, the cell
    given by its bodies, centre of mass com, mass M, half size hl and second
    moments Q (xx, yy, zz, xy, xz, yz), all in the cell's frame F (F maps the
    unit offset com - g to +z): inputs hl / d and F Q F^T / (M hl^2); targets
    F res / (M hl^2 / d^4) and F J F^T / (M hl^2 / d^5) (xx, yy, xy, xz, yz)."""
    src = np.stack([bx, by, bz], axis=1)
    a, G = field_and_gradient(src, bm, g, eps2)
    am, Gm = field_and_gradient(np.asarray(com)[None, :], [M], g, eps2)
    res, dres = a - am, G - Gm
    rvec = np.asarray(com, dtype=np.float64) - np.asarray(g, dtype=np.float64)
    d = np.sqrt(rvec @ rvec)
    F = frame(rvec / d)
    qp = F @ sym_matrix(np.asarray(Q, dtype=np.float64) / (M * hl * hl)) @ F.T
    x = np.zeros(IN)
    x[0] = hl / d
    x[1:7] = [qp[0, 0], qp[1, 1], qp[2, 2], qp[0, 1], qp[0, 2], qp[1, 2]]
    sres = M * hl * hl / d ** 4
    sgrad = sres / d
    rp = F @ res / sres
    jp = F @ dres @ F.T / sgrad
    t = np.array([rp[0], rp[1], rp[2], jp[0, 0], jp[1, 1], jp[0, 1], jp[0, 2], jp[1, 2]])
    return x, t


def correction_from_outputs(y, com, M, hl, g):
    """Residual (3) and gradient (3x3) at g in the original axes from the 8
    used outputs of one pair (the traversal's back-rotation)."""
    rvec = np.asarray(com, dtype=np.float64) - np.asarray(g, dtype=np.float64)
    d = np.sqrt(rvec @ rvec)
    F = frame(rvec / d)
    sres = M * hl * hl / d ** 4
    sgrad = sres / d
    jp = np.array([[y[3], y[5], y[6]], [y[5], y[4], y[7]], [y[6], y[7], -y[3] - y[4]]])
    return F.T @ np.asarray(y[:3]) * sres, F.T @ jp @ F * sgrad


def second_moments(bx, by, bz, bm):
    """Q about the centre of mass: xx, yy, zz, xy, xz, yz (float64)."""
    m = np.asarray(bm, dtype=np.float64)
    M = m.sum()
    c = np.array([np.dot(m, bx), np.dot(m, by), np.dot(m, bz)]) / M
    dx, dy, dz = bx - c[0], by - c[1], bz - c[2]
    return np.array([np.dot(m, dx * dx), np.dot(m, dy * dy), np.dot(m, dz * dz),
                     np.dot(m, dx * dy), np.dot(m, dx * dz), np.dot(m, dy * dz)])


def quadrupole_field(Q, r):
    """Closed-form quadrupole field at offset -r from the centre of mass of a
    mass distribution with second moments Q (xx, yy, zz, xy, xz, yz): the
    leading term of the monopole residual, a_i = (3 / 2) [5 (r.T.r) r_i / d^7
    - 2 (T r)_i / d^5 - tr(T) r_i / d^5] with T = Q, r from the field point to
    the centre of mass (unsoftened)."""
    T = np.array([[Q[0], Q[3], Q[4]], [Q[3], Q[1], Q[5]], [Q[4], Q[5], Q[2]]])
    r = np.asarray(r, dtype=np.float64)
    d = np.sqrt(r @ r)
    return 1.5 * (5.0 * (r @ T @ r) * r / d ** 7 - 2.0 * (T @ r) / d ** 5
                  - np.trace(T) * r / d ** 5)
