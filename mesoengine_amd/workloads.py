"""The five workloads of the voxel hot path, with their seeds and cameras.

Cameras are float32 values computed here once, so the GPU path and any scalar
implementation start from the same bits.  Up vectors: the dense sphere grid
uses +Z up like FVoxelCamera (Runtimes/Instance/VoxelCamera.h:23); terrain
scenes use +Y up, the axis the reference's terrain generator treats as height
(Runtimes/Helper/GeneratorHelper.h:104).
"""
import math

import numpy as np

LIGHT = (0.4, 1.0, 0.3)  # toward the light; normalised by the library


def _f32(v):
    return [float(np.float32(x)) for x in v]


# ---- dense 128^3 u8, 8 spheres, 256x256, camera at 1.5x diagonal
SPHERES = dict(n=128, seed=1, n_spheres=8, width=256, height=256, vfov=60.0)


def spheres_camera(n=128, width=256, height=256, direction=(1.0, 0.8, 0.6)):
    c = np.array([n / 2] * 3, np.float64)
    d = np.array(direction, np.float64)
    d /= np.linalg.norm(d)
    eye = c + 1.5 * n * math.sqrt(3.0) * d
    return dict(eye=_f32(eye), target=_f32(c), up=[0.0, 0.0, 1.0], vfov=60.0, width=width,
                height=height)


# ---- 512^3 brick map heightfield, 1920x1080 primary + shadow
TERRAIN = dict(n=512, seed=2, width=1920, height=1080, vfov=60.0)


def terrain_camera(n=512, width=1920, height=1080):
    s = n / 512.0
    return dict(eye=_f32([-60 * s, 260 * s, 256 * s]), target=_f32([300 * s, 30 * s, 256 * s]),
                up=[0.0, 1.0, 0.0], vfov=60.0, width=width, height=height)


# ---- 4096 chunks of 32^3 u16
CHUNKS = dict(n_chunks=4096, seed=3)

# ---- serving over 16 resident terrain/cave scenes
SERVE_RESOLUTIONS = [(256, 256), (640, 360), (1280, 720), (1920, 1080)]


def serve_scenes(seed=4, count=16):
    rng = np.random.default_rng(seed)
    edges = []
    for _ in range(count):
        u = rng.uniform(math.log(256), math.log(1024))
        edges.append(int(8 * round(math.exp(u) / 8)))
    return edges


def serve_requests(n_scenes_edges, n_requests, qps, seed=4):
    """Poisson arrivals; uniform resolutions; cameras on a sphere of radius
    1.5 * L * sqrt(3) around each scene's centre."""
    rng = np.random.default_rng(seed + 1000)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / qps)
        s = int(rng.integers(len(n_scenes_edges)))
        L = n_scenes_edges[s]
        w, h = SERVE_RESOLUTIONS[int(rng.integers(len(SERVE_RESOLUTIONS)))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        v[1] = abs(v[1]) * 0.8 + 0.2  # above the terrain
        v /= np.linalg.norm(v)
        c = np.array([L / 2, L * 0.15, L / 2])
        eye = c + 1.5 * L * math.sqrt(3.0) * v
        reqs.append(dict(arrival=t, scene=s, width=w, height=h, eye=_f32(eye), target=_f32(c),
                         up=[0.0, 1.0, 0.0], vfov=60.0))
    return reqs


# ---- 2048^3 two-level caves, 3840x2160, primary + shadow + 4 AO
CAVES = dict(n=2048, seed=5, width=3840, height=2160, ao_samples=4, ao_radius=16.0)


def caves_camera(n=2048, width=3840, height=2160):
    return dict(eye=_f32([n * 0.5 + 0.3, n * 0.5 + 0.2, n * 0.5 + 0.1]),
                target=_f32([n * 0.9, n * 0.45, n * 0.8]), up=[0.0, 1.0, 0.0], vfov=60.0,
                width=width, height=height)
