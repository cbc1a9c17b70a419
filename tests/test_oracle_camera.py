"""The oracle's pinhole camera against the reference's projection, in FP64.

The reference projects with glm::perspective(Fov, Width/Height, ...), a
vertical field of view (Runtimes/Instance/VoxelCamera.cpp:12-17), default
60 degrees (Runtimes/Instance/VoxelWindowsInstance.h:22), and a right-handed
glm::lookAt view whose forward is -row 2 of the view matrix
(VoxelCamera.cpp:24-29).  lightweightvk binds its viewports with a negative
height (ThirdParty/lightweightvk/lvk/vulkan/VulkanClasses.cpp:2350-2361), so
NDC +y (camera up) lands on framebuffer row 0.

Here the view and projection matrices are built independently, as glm
builds them, in float64; world points on the oracle's primary rays are
projected through them and must land on the centre of the pixel that cast
the ray.  A horizontal-fov camera, a half-angle fov, a vertically or
horizontally flipped image, or a left-handed basis each moves those points
by many pixels.
"""
import math

import numpy as np
import pytest

import oracle as O


def f32(v):
    """the camera API takes float32 vectors: the float64 reference starts
    from the same rounded values"""
    return np.asarray(v, np.float32).astype(np.float64)


def look_at(eye, center, up):
    """glm::lookAt (right-handed): rows s, u, -f and the eye translation."""
    eye, center, up = (f32(v) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def perspective(fovy_rad, aspect, near, far):
    """glm::perspective, right-handed, depth 0..1 (GLM_FORCE_DEPTH_ZERO_TO_ONE)."""
    t = math.tan(fovy_rad / 2.0)
    p = np.zeros((4, 4))
    p[0, 0] = 1.0 / (aspect * t)
    p[1, 1] = 1.0 / t
    p[2, 2] = far / (near - far)
    p[2, 3] = -(far * near) / (far - near)
    p[3, 2] = -1.0
    return p


def to_pixel(world, view, proj, w, h):
    """World point -> framebuffer coordinates through lvk's flipped viewport
    (y_fb = height - (ndc_y + 1) * height / 2)."""
    c = proj @ view @ np.append(np.asarray(world, np.float64), 1.0)
    ndc = c[:3] / c[3]
    return (ndc[0] + 1.0) * 0.5 * w, h - (ndc[1] + 1.0) * 0.5 * h, c[3]


CAMERAS = [
    # eye, target, up, vfov, width, height
    ([300.0, 250.0, 200.0], [64.0, 64.0, 64.0], [0.0, 0.0, 1.0], 60.0, 256, 256),
    ([-60.0, 260.0, 256.0], [300.0, 30.0, 256.0], [0.0, 1.0, 0.0], 60.0, 1920, 1080),
    ([1024.3, 1024.2, 1024.1], [1843.2, 921.6, 1638.4], [0.0, 1.0, 0.0], 60.0, 3840, 2160),
    ([10.0, -20.0, 5.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 35.0, 333, 121),
    ([5.0, 5.0, 40.0], [6.0, 4.0, 0.0], [0.0, 1.0, 0.0], 90.0, 97, 203),
]


def _pixels(w, h):
    pts = [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1), (w // 2, h // 2), (w // 3, 2 * h // 3)]
    rng = np.random.default_rng(w * 7 + h)
    pts += [(int(rng.integers(w)), int(rng.integers(h))) for _ in range(24)]
    return pts


@pytest.mark.parametrize("cam", CAMERAS)
def test_primary_rays_project_to_their_pixel_centres(cam):
    eye, target, up, vfov, w, h = cam
    c = O.camera(eye, target, up, vfov, w, h)
    view = look_at(eye, target, up)
    proj = perspective(math.radians(vfov), w / h, 0.1, 1e5)
    for px, py in _pixels(w, h):
        o, d = O.primary_ray(c, px, py)
        assert np.allclose(o, eye)
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1.0) < 1e-6
        for dist in (1.0, 37.5, 900.0):
            p = o.astype(np.float64) + dist * d.astype(np.float64)
            x, y, depth = to_pixel(p, view, proj, w, h)
            assert depth > 0.0, "point behind the camera"
            # float32 ray setup: well under a hundredth of a pixel
            assert abs(x - (px + 0.5)) < 5e-3 * max(1.0, w / 1000), (cam, px, py, x)
            assert abs(y - (py + 0.5)) < 5e-3 * max(1.0, h / 1000), (cam, px, py, y)


@pytest.mark.parametrize("cam", CAMERAS)
def test_fov_is_vertical_and_row_zero_is_up(cam):
    eye, target, up, vfov, w, h = cam
    c = O.camera(eye, target, up, vfov, w, h)
    fwd = f32(target) - f32(eye)
    fwd /= np.linalg.norm(fwd)
    th = math.tan(math.radians(vfov) / 2.0)

    def ray(px, py):
        return O.primary_ray(c, px, py)[1].astype(np.float64)

    def angle(a, b):
        return math.atan2(float(np.linalg.norm(np.cross(a, b))), float(a @ b))

    # top and bottom rows at a centre column subtend 2 atan(tan(vfov/2) (h-1)/h)
    # when w is odd (the column's rays lie in the vertical plane)
    if w % 2 == 1:
        col = w // 2
        top, bot = ray(col, 0), ray(col, h - 1)
        assert abs(angle(top, bot) - 2 * math.atan(th * (h - 1) / h)) < 2e-6
        # row 0 is on the camera-up side
        upv = np.cross(np.cross(fwd, f32(up)), fwd)
        assert top @ upv > 0.0 > bot @ upv
    # left and right columns at a centre row: horizontal fov from the aspect
    if h % 2 == 1:
        row = h // 2
        left, right = ray(0, row), ray(w - 1, row)
        assert abs(angle(left, right) - 2 * math.atan(th * (w / h) * (w - 1) / w)) < 2e-6
        # column 0 is on the left: right = normalize(fwd x up)
        r = np.cross(fwd, f32(up))
        assert left @ r < 0.0 < right @ r
    # odd sizes: the centre pixel looks at the target
    if w % 2 == 1 and h % 2 == 1:
        assert angle(ray(w // 2, h // 2), fwd) < 1e-6


def test_basis_is_right_handed_lookat():
    eye, target, up = [1.0, 2.0, 3.0], [4.0, -2.0, 0.5], [0.0, 0.0, 1.0]
    c = O.camera(eye, target, up, 60.0, 64, 48)
    f = np.asarray(target) - np.asarray(eye)
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    assert np.allclose(np.asarray(c.fwd, np.float64), f, atol=1e-6)
    assert np.allclose(np.asarray(c.right, np.float64), s, atol=1e-6)
    assert np.allclose(np.asarray(c.up, np.float64), u, atol=1e-6)
    # the half-extents of the image plane at unit distance
    assert abs(c.sy - math.tan(math.radians(30.0))) < 1e-6
    assert abs(c.sx - math.tan(math.radians(30.0)) * 64 / 48) < 1e-6
