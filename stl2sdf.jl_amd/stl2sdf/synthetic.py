"""Seeded synthetic closed meshes used by the tests and the benchmark.

* icosphere(r, subdiv): icosahedron subdivided ``subdiv`` times, vertices
  projected to the sphere of radius r (subdiv=4: 5120 triangles).
* star_surface(n_theta, n_phi, seed): r(theta, phi) = 1 + 0.3 * sum of 8 real
  orthonormal spherical harmonics (l <= 6, l and m uniform, N(0,1)
  coefficients) on an n_theta x n_phi latitude-longitude mesh with pole caps.
  The radius is rescaled if its minimum would drop below 0.2.
* shells(...): small disconnected icospheres scattered in a box.

Meshes are returned as (X, IEN): X 3 x nnp float64, IEN 3 x nel int64
(1-based), both column-major, outward orientation (counter-clockwise seen
from outside).  ``to_float32`` rounds vertices through float32 as a binary
STL round trip would.
"""
import numpy as np
from scipy.special import sph_harm_y


def _ico_base():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
                 dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(radius=1.0, subdiv=4, center=(0.0, 0.0, 0.0)):
    v, f = _ico_base()
    verts = list(v)
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, dtype=np.int64)
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    return np.asfortranarray(V.T), np.asfortranarray((f + 1).T)


def _real_sph(l, m, theta, phi):
    # real orthonormal spherical harmonics from scipy's complex ones
    if m > 0:
        return np.sqrt(2.0) * (-1) ** m * np.real(sph_harm_y(l, m, theta, phi))
    if m < 0:
        return np.sqrt(2.0) * (-1) ** m * np.imag(sph_harm_y(l, -m, theta, phi))
    return np.real(sph_harm_y(l, 0, theta, phi))


def star_radius_fn(seed=0, n_terms=8, lmax=6, amp=0.3):
    rng = np.random.Generator(np.random.MT19937(seed))
    terms = []
    for _ in range(n_terms):
        l = int(rng.integers(0, lmax + 1))
        m = int(rng.integers(-l, l + 1))
        c = float(rng.standard_normal())
        terms.append((l, m, c))

    def r(theta, phi):
        s = np.zeros(np.broadcast(theta, phi).shape)
        for l, m, c in terms:
            s = s + c * _real_sph(l, m, theta, phi)
        return 1.0 + amp * s

    return r, terms


def star_surface(n_theta=450, n_phi=450, seed=0):
    """Closed star-shaped surface on a lat-long mesh: (X, IEN, info)."""
    rfn, terms = star_radius_fn(seed)
    th = np.pi * np.arange(1, n_theta) / n_theta       # interior rings
    ph = 2.0 * np.pi * np.arange(n_phi) / n_phi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    R = rfn(TH, PH)
    r_np, r_sp = float(rfn(0.0, 0.0)), float(rfn(np.pi, 0.0))
    rmin = min(R.min(), r_np, r_sp)
    scale = 1.0
    if rmin < 0.2:
        # shrink the perturbation so the body stays star-shaped
        scale = 0.8 / (1.0 - rmin)
        R = 1.0 + (R - 1.0) * scale
        r_np = 1.0 + (r_np - 1.0) * scale
        r_sp = 1.0 + (r_sp - 1.0) * scale
    P = np.stack([R * np.sin(TH) * np.cos(PH), R * np.sin(TH) * np.sin(PH),
                  R * np.cos(TH)], axis=-1).reshape(-1, 3)
    north = np.array([[0.0, 0.0, r_np]])
    south = np.array([[0.0, 0.0, -r_sp]])
    V = np.concatenate([north, P, south])
    nr = n_theta - 1
    idx = 1 + np.arange(nr * n_phi).reshape(nr, n_phi)
    s_id = 1 + nr * n_phi
    faces = []
    j = np.arange(n_phi)
    jn = (j + 1) % n_phi
    faces.append(np.stack([np.zeros(n_phi, np.int64), idx[0, j], idx[0, jn]], 1))
    for i in range(nr - 1):
        a, b, c, d = idx[i, j], idx[i + 1, j], idx[i + 1, jn], idx[i, jn]
        faces.append(np.stack([a, b, c], 1))
        faces.append(np.stack([a, c, d], 1))
    faces.append(np.stack([np.full(n_phi, s_id), idx[nr - 1, jn], idx[nr - 1, j]], 1))
    F = np.concatenate(faces)
    info = {"terms": terms, "scale": scale, "nel": int(F.shape[0]), "nnp": int(V.shape[0])}
    return np.asfortranarray(V.T), np.asfortranarray((F + 1).T), info


def shells(n=40, rmin=0.01, rmax=0.05, box_min=(-1, -1, -1), box_max=(1, 1, 1), seed=1):
    """n icospheres subdivided once (80 triangles each) at random centres."""
    rng = np.random.Generator(np.random.MT19937(seed))
    Xs, Is, off = [], [], 0
    lo, hi = np.asarray(box_min, float), np.asarray(box_max, float)
    for _ in range(n):
        r = rng.uniform(rmin, rmax)
        c = lo + (hi - lo) * rng.uniform(size=3)
        X, IEN = icosphere(r, 1, c)
        Xs.append(X)
        Is.append(IEN + off)
        off += X.shape[1]
    return np.asfortranarray(np.concatenate(Xs, 1)), np.asfortranarray(np.concatenate(Is, 1))


def merge(*meshes):
    Xs, Is, off = [], [], 0
    for X, IEN in meshes:
        Xs.append(X)
        Is.append(IEN + off)
        off += X.shape[1]
    return np.asfortranarray(np.concatenate(Xs, 1)), np.asfortranarray(np.concatenate(Is, 1))


def to_float32(X):
    return np.asfortranarray(np.asarray(X, np.float32).astype(np.float64))


def write_stl_binary(path, X, IEN):
    """Binary STL (80-byte header, uint32 count, 50-byte records)."""
    X = np.asarray(X, np.float64)
    IEN = np.asarray(IEN, np.int64)
    tri = X[:, IEN - 1]                      # 3 x 3 x nel
    v = np.transpose(tri, (2, 1, 0))         # nel x vertex x xyz
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    nn = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(nn > 0, n / np.where(nn > 0, nn, 1), 0)
    rec = np.zeros(len(v), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    rec["n"] = n
    rec["v"] = v
    with open(path, "wb") as f:
        f.write(b"binary stl written by stl2sdf synthetic".ljust(80, b" "))
        f.write(np.uint32(len(v)).tobytes())
        f.write(rec.tobytes())


def batch_family(count=64, seed=2, tri_range=(2000, 500000), grid_points=(64, 96, 128, 192)):
    """Seeded batch of star surfaces: triangle counts log-uniform in
    tri_range, each with its own harmonic seed and a grid with a uniformly
    chosen number of points on its longest axis.  Returns a list of
    (X, IEN, points_on_longest_axis, info) with float32-rounded vertices."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(count):
        tris = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        n = max(4, int(round((1.0 + np.sqrt(1.0 + 2.0 * tris)) / 2.0)))
        g = int(rng.choice(grid_points))
        X, IEN, info = star_surface(n, n, seed=int(rng.integers(0, 2 ** 31)))
        info["target_tris"] = tris
        out.append((to_float32(X), IEN, g, info))
    return out
