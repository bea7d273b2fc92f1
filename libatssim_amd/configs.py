"""Named workloads of the VBD path (synthetic, seeded).

grid_c0      32x32 grid (1,024 vertices), 1 cm spacing, z-jitter U(-1e-4, 1e-4),
             two top corners pinned, m = 1e-3 kg, k_stretch 1e4, k_bend 1e-2,
             h = 1/60 s, 100 iterations/frame, 30 frames.
grid_c1      256x256 grid (65,536 vertices), same material; top row pinned;
             per-frame wind of total magnitude U(0, 2) N, direction U(0, 2 pi)
             in the xz plane, split evenly over the vertices; 100 iterations,
             60 frames.
batch_c2     512 independent grid cloths, side n ~ U{16..128} (~3.2 M vertices),
             k_stretch log-U[1e3, 1e5], k_bend log-U[1e-3, 1e-1] per cloth,
             1-4 random boundary vertices pinned, m = 1e-3 kg, 100 iterations;
             all draws from splitmix64(seed).
sphere_c4    (sphere_scene) 1024x1024 horizontal grid (2 mm spacing, 2.05 m wide, no pins)
             10 cm above a fixed sphere of radius 0.3 m (icosphere, 20,480
             faces); contact thickness 2 mm, k_contact 1e5, detection radius
             2.5 cm (one frame of fall at impact speed), hash cell 2x the
             edge length (4 mm; every face is entered in each cell its
             bounding box grown by the radius touches, so one cell lookup
             finds every face within the radius);
             50 iterations/frame, 120 frames.  Warm-start training drops
             (sphere_train_scenario): 384x384 cloths of the same spacing,
             dropped from 6-14 cm, centre offset up to +-10 cm.
warm_c3      learned warm start on the grid_c1 cloth: training scenarios k =
             0..31 of 200 frames each, per-scenario wind seed and pin set
             (k % 4: 0, 2 top row; 1 two top corners; 3 one to four random
             top-boundary vertices plus both top corners); evaluation on the
             grid_c1 scene itself (top row pinned, its wind).
square8k     the reference's example-2 run (examples/example2_async_iter.cpp:
             133-392, 715-749, 1713-1810): square8K.obj, pins where the
             AABB-normalised z < 0.01 and x < 0.01 or x > 0.99, m = 0.01 kg,
             k_stretch 1e4, k_bend 5e-3 * 0.5, 100 iterations, 30 frames.
"""
import numpy as np

from .vbd import Cloth, Params, make_grid, make_icosphere, make_wind

SEED_C0 = 20250101
SEED_C1 = 20250102
SEED_WIND = 20250103
SEED_C2 = 20250104
SEED_C3 = 20250106


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): the same stream as the C++ scenes."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def batch_c2_spec(num_cloths=512, seed=SEED_C2, nmin=16, nmax=128):
    """Per cloth: (side n, k_stretch, k_bend, pinned vertex ids, grid seed)."""
    rng = SplitMix64(seed)
    spec = []
    for k in range(num_cloths):
        n = nmin + min(int(rng.uniform() * (nmax - nmin + 1)), nmax - nmin)
        ks = 10.0 ** (3.0 + 2.0 * rng.uniform())
        kb = 10.0 ** (-3.0 + 2.0 * rng.uniform())
        npins = 1 + min(int(rng.uniform() * 4), 3)
        boundary = [c for c in range(n)] + [(n - 1) * n + c for c in range(n)] + \
                   [r * n for r in range(1, n - 1)] + [r * n + n - 1 for r in range(1, n - 1)]
        pins = []
        while len(pins) < npins:
            b = boundary[min(int(rng.uniform() * len(boundary)), len(boundary) - 1)]
            if b not in pins:
                pins.append(b)
        spec.append((n, ks, kb, sorted(pins), seed * 1000003 + k))
    return spec


def batch_c2(num_cloths=512, seed=SEED_C2, nmin=16, nmax=128):
    """(cloths, per-cloth k_stretch, per-cloth k_bend, params)."""
    cloths, ks, kb = [], [], []
    for n, s, b, pins, gseed in batch_c2_spec(num_cloths, seed, nmin, nmax):
        x, f = make_grid(n, n, 0.01, 1e-4, gseed)
        fixed = np.zeros(n * n, np.uint8)
        fixed[pins] = 1
        cloths.append(Cloth(x, f, np.float32(1e-3), fixed))
        ks.append(s)
        kb.append(b)
    return cloths, np.array(ks, np.float32), np.array(kb, np.float32), Params(constraint_iter_count=100)


def grid_cloth(n, pin="corners", seed=SEED_C0, spacing=0.01, jitter=1e-4, mass=1e-3):
    x, f = make_grid(n, n, spacing, jitter, seed)
    fixed = np.zeros(n * n, np.uint8)
    if pin == "corners":
        fixed[0] = 1
        fixed[n - 1] = 1
    elif pin == "top":
        fixed[:n] = 1
    return Cloth(x, f, np.float32(mass), fixed)


def grid_c0():
    return grid_cloth(32, "corners", SEED_C0), Params(constraint_iter_count=100), 30, None


def grid_c1(frames=60):
    cloth = grid_cloth(256, "top", SEED_C1)
    wind = make_wind(frames, cloth.num_verts, 2.0, SEED_WIND)
    return cloth, Params(constraint_iter_count=100), frames, wind


def warm_c3_scenario(k, n=256, frames=200, seed=SEED_C3):
    """Training scenario k of the learned warm start: (cloth, wind[frames])."""
    rng = SplitMix64(seed * 1000003 + k)
    fixed = np.zeros(n * n, np.uint8)
    kind = k % 4
    if kind in (0, 2):
        fixed[:n] = 1
    else:
        fixed[0] = fixed[n - 1] = 1
        if kind == 3:
            for _ in range(1 + min(int(rng.uniform() * 4), 3)):
                fixed[min(int(rng.uniform() * n), n - 1)] = 1
    x, f = make_grid(n, n, 0.01, 1e-4, SEED_C1)
    wind = make_wind(frames, n * n, 2.0, rng.next() & ((1 << 63) - 1))
    return Cloth(x, f, np.float32(1e-3), fixed), wind


def horizontal_grid(n, spacing, height, seed, jitter=1e-4):
    """n x n grid in the xz plane at y = height (grid rows along -z)."""
    x, f = make_grid(n, n, spacing, jitter, seed)
    xh = np.empty_like(x)
    xh[:, 0] = x[:, 0]
    xh[:, 1] = np.float32(height) + x[:, 2]
    xh[:, 2] = -x[:, 1] - np.float32(0.5 * (n - 1) * spacing)
    return xh.astype(np.float32), f


def sphere_scene(n=1024, spacing=0.002, radius=0.3, drop=0.1, level=5, seed=20250105, iters=50):
    """(cloth, params, obstacle (x, faces), contact parameters dict)."""
    x, f = horizontal_grid(n, spacing, radius + drop, seed)
    cloth = Cloth(x, f, np.float32(1e-3), np.zeros(n * n, np.uint8))
    ox, of = make_icosphere(level, radius)
    contact = dict(thickness=0.002, stiffness=1e5, radius=0.025, cell=2.0 * spacing)
    return cloth, Params(constraint_iter_count=iters), (ox, of), contact


SEED_C4T = 20250107


def sphere_train_scenario(k, n=384, spacing=0.002, radius=0.3, seed=SEED_C4T):
    """Training drop k of the sphere regime: a smaller cloth of the same
    spacing, material and contact model, dropped from 6-14 cm with its centre
    offset by up to +-10 cm from the sphere's top.  No wind."""
    rng = SplitMix64(seed * 1000003 + k)
    drop = 0.06 + 0.08 * rng.uniform()
    ox = (rng.uniform() - 0.5) * 0.2
    oz = (rng.uniform() - 0.5) * 0.2
    x, f = horizontal_grid(n, spacing, radius + drop, seed + k)
    x[:, 0] += np.float32(ox)
    x[:, 2] += np.float32(oz)
    return Cloth(x, f, np.float32(1e-3), np.zeros(n * n, np.uint8)), None


def square8k_pins(x):
    mn = x.min(0)
    rng = (x.max(0) - mn).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = (np.float32(1) / rng).astype(np.float32)
        nrm = ((x - mn) * inv).astype(np.float32)
    return ((nrm[:, 2] < np.float32(0.01)) & ((nrm[:, 0] > np.float32(0.99)) | (nrm[:, 0] < np.float32(0.01)))).astype(np.uint8)


def square8k_params():
    return Params(implicit_dt=np.float32(1) / np.float32(60), constraint_iter_count=100,
                  stiffness_stretch_spring=1e4,
                  stiffness_quadratic_bending=float(np.float32(5e-3) * np.float32(0.5)), damping_cloth=2.0)


def square8k_cloth(x, faces):
    return Cloth(x, faces, np.float32(0.01) * np.float32(1.0), square8k_pins(x))
