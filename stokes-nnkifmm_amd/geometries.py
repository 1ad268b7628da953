"""Host-side point generators for the benchmark configurations.

These are this library's own generators (the reference's Geometries/ are
input scripts outside the accelerated path):
  cube_sphere       6 n^2 face-centred cube points projected to a sphere
                    (384 force points at n = 8, 6144 quadrature points at 32)
  prolate_spheroid  the same construction stretched along z (semi-axes
                    a, a, c)
  plane_wall        an n x n grid on a square in the plane z = z0
  random_frames     seeded non-overlapping centres and random orientations
  spheroid_suspension  bodies + wall assembled into a Swimmers system with
                    NEAREST quadrature (the adaptive mobility workload:
                    64 prolate spheroids, aspect 3, semi-major axis 1, in a
                    20^3 box above a 40 x 40 wall, 4x refined quadrature)
  squirmer_slip     tangential squirming velocity B1 sin(t) + B2 sin(t) cos(t)
  squirmer_pair     two spherical squirmers (force- and torque-free)
  batch_bodies      independent single-body systems (spheres and
                    spheroids, 200-5000 force points log-uniform, 4x
                    refined quadrature, eps uniform in [0.002, 0.02])
  wall, prolateSpheroid, Rotate, unique_rows
                    the reference's generators (Geometries/wall.m,
                    prolateSpheroid.m, Rotate.m; unique(...,'rows','stable'))
  squirmer_box      the suspension of Data Generation/Squiremers.m: squirmers
                    in a closed box of six walls
"""
import numpy as np


def cube_sphere(n, radius=1.0, centre=(0.0, 0.0, 0.0)):
    t = (np.arange(n) + 0.5) / n * 2 - 1
    a, b = np.meshgrid(t, t, indexing="ij")
    a, b = a.ravel(), b.ravel()
    faces = []
    for axis in range(3):
        for s in (-1.0, 1.0):
            p = np.empty((len(a), 3))
            p[:, axis] = s
            p[:, (axis + 1) % 3] = a
            p[:, (axis + 2) % 3] = b
            faces.append(p)
    p = np.concatenate(faces)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return radius * p + np.asarray(centre, float)


def prolate_spheroid(n, a, c, centre=(0.0, 0.0, 0.0)):
    p = cube_sphere(n)
    return p * np.array([a, a, c]) + np.asarray(centre, float)


def plane_wall(n, half_width, z0=0.0, centre=(0.0, 0.0)):
    g = (np.arange(n) + 0.5) / n * 2 * half_width - half_width
    X, Y = np.meshgrid(g + centre[0], g + centre[1], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, float(z0))], 1)


def random_frames(nbody, box, min_dist, zmin, seed):
    """Centres uniform in [-box/2+1, box/2-1]^2 x [zmin, box-1], at least
    min_dist apart (rejection sampling), and uniformly random orientations
    (QR of a Gaussian matrix); returns (centres, b1, b2)."""
    rng = np.random.default_rng(seed)
    lo = np.array([-box / 2 + 1, -box / 2 + 1, zmin])
    hi = np.array([box / 2 - 1, box / 2 - 1, box - 1])
    centres = []
    tries = 0
    while len(centres) < nbody:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("random_frames: cannot place the bodies")
        c = rng.uniform(lo, hi)
        if all(np.linalg.norm(c - d) >= min_dist for d in centres):
            centres.append(c)
    b1, b2 = [], []
    for _ in range(nbody):
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 2] = -q[:, 2]
        b1.append(q[:, 0])
        b2.append(q[:, 1])
    return np.array(centres), np.array(b1), np.array(b2)


def spheroid_suspension(nbody=64, n=18, refine=4, a=1.0 / 3.0, c=1.0, box=20.0, wall_n=40, seed=7,
                        force=(0.0, 0.0, -1.0)):
    """nbody prolate spheroids (semi-axes a, a, c; 6 n^2 force points,
    6 (refine n)^2 quadrature points with the nearest map) at random
    non-overlapping positions and orientations in a box^3 region above a
    wall_n x wall_n no-slip wall at z = 0 spanning the box (quadrature
    refine x finer).  Each body carries the force `force` and no moment.
    Returns a Swimmers system (tree and operator not built yet)."""
    from .nearest import nearestMatrix
    from .swimmers import Swimmers
    body = prolate_spheroid(n, a, c)
    fine = prolate_spheroid(refine * n, a, c)
    NN = nearestMatrix(body, fine)
    centres, b1, b2 = random_frames(nbody, box, 2.0 * c + 0.1, c + 0.2, seed)
    sw = Swimmers()
    for i in range(nbody):
        sw.addSwimmer(body, fine, NN, centres[i], np.concatenate([b1[i], b2[i]]), np.zeros_like(body),
                      np.asarray(force, float), np.zeros(3))
    if wall_n > 0:
        wall = plane_wall(wall_n, box / 2)
        wfine = plane_wall(refine * wall_n, box / 2)
        sw.updateBnd(wall, wfine, nearestMatrix(wall, wfine), np.zeros_like(wall))
    return sw


def batch_bodies(nprob=256, nmin=200, nmax=5000, refine=4, eps_range=(0.002, 0.02), seed=11):
    """nprob single-body Swimmers systems for the batched mobility solve:
    a sphere (radius 1) or a prolate spheroid (semi-axes 1/3, 1/3, 1) with
    6 n^2 force points, n from a log-uniform draw of the point count in
    [nmin, nmax], and 6 (refine n)^2 quadrature points; random orientation,
    a unit force in a random direction and a random moment of size 0.1.
    Returns (systems, eps)."""
    from .nearest import nearestMatrix
    from .swimmers import Swimmers
    rng = np.random.default_rng(seed)
    out, eps, cache = [], [], {}
    for _ in range(nprob):
        npts = np.exp(rng.uniform(np.log(nmin), np.log(nmax)))
        n = max(2, int(round(np.sqrt(npts / 6.0))))
        kind = int(rng.integers(2))
        key = (kind, n)
        if key not in cache:
            body = cube_sphere(n) if kind == 0 else prolate_spheroid(n, 1.0 / 3.0, 1.0)
            fine = cube_sphere(refine * n) if kind == 0 else prolate_spheroid(refine * n, 1.0 / 3.0, 1.0)
            cache[key] = (body, fine, nearestMatrix(body, fine))
        body, fine, NN = cache[key]
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 2] = -q[:, 2]
        F = rng.standard_normal(3)
        F /= np.linalg.norm(F)
        M = rng.standard_normal(3)
        M *= 0.1 / np.linalg.norm(M)
        sw = Swimmers()
        sw.addSwimmer(body, fine, NN, rng.uniform(-5, 5, 3), np.concatenate([q[:, 0], q[:, 1]]),
                      np.zeros_like(body), F, M)
        out.append(sw)
        eps.append(rng.uniform(*eps_range))
    return out, np.array(eps)


def squirmer_slip(points, B1=1.0, B2=0.0, axis=(0.0, 0.0, 1.0)):
    """Surface slip of a spherical squirmer in its body frame:
    u = (B1 sin t + B2 sin t cos t) e_t, t the polar angle from `axis`,
    e_t = (cos t r - e) / sin t the polar unit vector (zero at the poles).
    A lone squirmer then swims along `axis` at 2 B1 / 3."""
    e = np.asarray(axis, float)
    e = e / np.linalg.norm(e)
    r = np.asarray(points, float)
    r = r / np.linalg.norm(r, axis=1, keepdims=True)
    ct = r @ e
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    et = np.zeros_like(r)
    ok = st > 1e-12
    et[ok] = (ct[ok, None] * r[ok] - e) / st[ok, None]
    return ((B1 * st + B2 * st * ct)[:, None]) * et


def squirmer_pair(n=16, refine=4, gap=0.5, B1=1.0, B2=0.0, tilt=0.5):
    """Two unit spheres (6 n^2 force points, 6 (refine n)^2 quadrature
    points, nearest map) a surface gap `gap` apart along x, swimming axes in
    the x-z plane at +/- tilt radians from z; force- and torque-free.
    Returns a Swimmers system (tree and operator not built)."""
    from .nearest import nearestMatrix
    from .swimmers import Swimmers
    body, fine = cube_sphere(n), cube_sphere(refine * n)
    NN = nearestMatrix(body, fine)
    vel = squirmer_slip(body, B1, B2)
    sw = Swimmers()
    d = 1.0 + gap / 2.0
    for sgn in (-1.0, 1.0):
        a = sgn * tilt
        b1 = [np.cos(a), 0.0, -np.sin(a)]   # body x axis
        b2 = [0.0, 1.0, 0.0]                # body y axis; body z = b1 x b2
        sw.addSwimmer(body, fine, NN, [sgn * d, 0.0, 0.0], b1 + b2, vel, np.zeros(3), np.zeros(3))
    return sw


def _linspace(a, b, n):
    """MATLAB linspace(a, b, n) (n >= 1; the last point is b)."""
    n = int(n)
    if n <= 1:
        return np.array([float(b)])
    return np.linspace(float(a), float(b), n)


def wall(centre, scale, axis, h):
    """Geometries/wall.m: a ceil(scale/h) grid on the plane through centre
    normal to `axis` (1 = x, 2 = y, 3 = z), in meshgrid column-major order."""
    c = np.asarray(centre, float).ravel()
    sc = np.asarray(scale, float).ravel()
    res = np.ceil(sc / h).astype(int)
    ax = int(axis)
    if ax not in (1, 2, 3):
        raise ValueError("wall: axis must be 1, 2 or 3")
    if ax == 1:
        u, v = (1, 2)
    elif ax == 2:
        u, v = (0, 2)
    else:
        u, v = (0, 1)
    gu = _linspace(c[u] - sc[0] / 2, c[u] + sc[0] / 2, res[0])
    gv = _linspace(c[v] - sc[1] / 2, c[v] + sc[1] / 2, res[1])
    U, V = np.meshgrid(gu, gv)                # rows follow gv, columns gu
    p = np.empty((U.size, 3))
    p[:, ax - 1] = c[ax - 1]
    p[:, u] = U.ravel(order="F")
    p[:, v] = V.ravel(order="F")
    return p


def prolateSpheroid(a, c, h):
    """Geometries/prolateSpheroid.m: the two poles, then rings of latitude
    v(i) = pi (i-1)/(vN-1) with ceil(2 pi a sqrt(1 - (z/c)^2) / h) points."""
    vN = int(np.ceil(np.pi * (3 * (a + c) - np.sqrt((3 * a + c) * (a + 3 * c))) / (2 * h)))
    v = _linspace(0, np.pi, vN)
    pts = [np.array([[0.0, 0.0, c], [0.0, 0.0, -c]])]
    for i in range(1, vN - 1):
        z = c * np.cos(v[i])
        uN = int(np.ceil(2 * np.pi * a * np.sqrt(1 - (z / c) ** 2) / h))
        u = _linspace(0, 2 * np.pi, uN + 1)[1:]
        pts.append(np.stack([a * np.cos(u) * np.sin(v[i]), a * np.sin(u) * np.sin(v[i]), np.full(uN, z)], 1))
    return np.concatenate(pts)


def Rotate(points, alpha, beta, gamma, centre):
    """Geometries/Rotate.m: rotate about centre by Rx(alpha) Ry(beta) Rz(gamma)
    (the matrices as written there)."""
    ca, sa, cb, sb, cg, sg = np.cos(alpha), np.sin(alpha), np.cos(beta), np.sin(beta), np.cos(gamma), np.sin(gamma)
    Rx = np.array([[1, 0, 0], [0, ca, sa], [0, -sa, ca]])
    Ry = np.array([[cb, 0, -sb], [0, 1, 0], [sb, 0, cb]])
    Rz = np.array([[cg, sg, 0], [-sg, cg, 0], [0, 0, 1]])
    c = np.asarray(centre, float).ravel()
    return ((Rx @ Ry @ Rz) @ (np.asarray(points, float) - c).T).T + c


def unique_rows(p):
    """unique(p, 'rows', 'stable'): first occurrence of each row, in order."""
    p = np.ascontiguousarray(p, float)
    _, first = np.unique(p, axis=0, return_index=True)
    return p[np.sort(first)]


def squirmer_box(nsw=10, seed=0, a=0.5, h_coarse=0.25, h_fine=0.13, box=10.0, wall_h=0.3, B1=2.0,
                 min_dist=1.5):
    """The suspension of Data Generation/Squiremers.m: `nsw` spherical
    squirmers (prolateSpheroid(a, a, h) bodies, coarse h_coarse, fine h_fine,
    nearest map, slip B1 sin(theta) e_theta about the body z axis) at seeded
    positions in [-4, 4]^3 at least min_dist apart with random orientations,
    inside a closed box of six box x box walls with spacing wall_h (the
    boundary is its own quadrature, NN = identity).  Returns (Swimmers,
    initial condition [x0; b1; b2] per body)."""
    from .nearest import nearestMatrix
    from .swimmers import Swimmers
    rng = np.random.default_rng(seed)
    hb = box / 2
    bnd = np.concatenate([wall([hb, 0, 0], [box, box], 1, wall_h), wall([-hb, 0, 0], [box, box], 1, wall_h),
                          wall([0, hb, 0], [box, box], 2, wall_h), wall([0, -hb, 0], [box, box], 2, wall_h),
                          wall([0, 0, hb], [box, box], 3, wall_h), wall([0, 0, -hb], [box, box], 3, wall_h)])
    bnd = unique_rows(bnd)
    sw = Swimmers()
    sw.updateBnd(bnd, bnd, nearestMatrix(bnd, bnd), np.zeros_like(bnd))
    coarse = prolateSpheroid(a, a, h_coarse)
    fine = prolateSpheroid(a, a, h_fine)
    NN = nearestMatrix(coarse, fine)
    vel = squirmer_slip(coarse, B1=B1)
    centres = []
    while len(centres) < nsw:
        x = -4 + 8 * rng.random(3)
        if all(np.linalg.norm(x - y) >= min_dist for y in centres):
            centres.append(x)
    y0 = []
    for x in centres:
        basis = Rotate(np.eye(3), *(2 * np.pi * rng.random(3)), [0, 0, 0])
        b1b2 = basis.ravel(order="F")[:6]
        sw.addSwimmer(coarse, fine, NN, x, b1b2, vel, [0, 0, 0], [0, 0, 0])
        y0.append(np.concatenate([x, b1b2]))
    return sw, np.concatenate(y0)
