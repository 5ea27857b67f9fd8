"""Deterministic synthetic scenes and cameras for the benchmark configurations.

Each builder returns a `Scene`: numpy arrays for meshes, instances, materials and lights,
plus an `hdr_scene_desc` view over them (the arrays are kept alive by the Scene object).
The same Scene feeds the MI355X library and the CPU checker, so both see byte-identical
inputs.

Configurations (sizes as named by the workload descriptions):
  * ``box_spheres``   256x256, closed 2 m box (12 triangles), two geodesic icospheres of
    frequency 11 (2 x 2420 triangles), one 0.5 x 0.5 m rect light, intensity 20, colour of a
    6500 K blackbody; UsdPreviewSurface materials with baseColor U[0.1,0.9]^3,
    roughness U[0.1,0.9], metallic in {0,1}.
  * ``many_lights``   1920x1080, 512x512-vertex height field (522,242 triangles) displaced by
    4-octave value noise + 500 instances of a 320-triangle UV sphere; 4096 lights (2048 point,
    2048 rect of side U[0.05,0.5] m), intensity 10^U[0,4], blackbody colour at U[2000,10000] K;
    a linear camera dolly.
"""
import ctypes as C
import math

import numpy as np

from . import abi

# ----------------------------------------------------------------------------- colour


def _cie_fits(lam):
    """Wyman et al. 2013 multi-lobe fits of the CIE 1931 matching functions (the same fits the
    renderer's spectral->RGB conversion uses)."""
    def g(x, mu, s1, s2):
        t = (x - mu) * np.where(x < mu, s1, s2)
        return np.exp(-0.5 * t * t)
    x = 0.362 * g(lam, 442.0, 0.0624, 0.0374) + 1.056 * g(lam, 599.8, 0.0264, 0.0323) \
        - 0.065 * g(lam, 501.1, 0.0490, 0.0382)
    y = 0.821 * g(lam, 568.8, 0.0213, 0.0247) + 0.286 * g(lam, 530.9, 0.0613, 0.0322)
    z = 1.217 * g(lam, 437.0, 0.0845, 0.0278) + 0.681 * g(lam, 459.0, 0.0385, 0.0725)
    return x, y, z


def blackbody_rgb(temp_k):
    """Normalised (max component = 1) linear-sRGB colour of a Planck spectrum at temp_k, white
    balanced to illuminant E like the renderer's SpectrumToRGB. Lights carry an RGB Color in
    the renderer, so a blackbody emitter is expressed as this colour."""
    lam = np.arange(360.0, 830.5, 1.0)
    lm = lam * 1e-9
    h, c, kb = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    b = (2 * h * c * c / lm ** 5) / np.expm1(h * c / (lm * kb * temp_k))
    xf, yf, zf = _cie_fits(lam)
    X, Y, Z = (b * xf).sum(), (b * yf).sum(), (b * zf).sum()
    r = 3.2404542 * X - 1.5371385 * Y - 0.4985314 * Z
    g = -0.9692660 * X + 1.8760108 * Y + 0.0415560 * Z
    bl = 0.0556434 * X - 0.2040259 * Y + 1.0572252 * Z
    rgb = np.maximum(np.array([r / 1.198185, g / 0.950366, bl / 0.907827]), 0.0)
    return (rgb / rgb.max()).astype(np.float32)


# ----------------------------------------------------------------------------- geometry


def box_mesh(half=1.0):
    """Closed axis-aligned cube [-half, half]^3, 12 triangles wound so the geometric normals
    point inward (the interior is the visible side)."""
    h = half
    p = np.array([[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h],
                  [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]], np.float32)
    # each quad (a,b,c,d) counter-clockwise seen from inside
    quads = [(0, 1, 2, 3),   # z = -h, normal +z
             (5, 4, 7, 6),   # z = +h, normal -z
             (4, 0, 3, 7),   # x = -h, normal +x
             (1, 5, 6, 2),   # x = +h, normal -x
             (4, 5, 1, 0),   # y = -h, normal +y
             (3, 2, 6, 7)]   # y = +h, normal -y
    idx = []
    for a, b, c, d in quads:
        idx += [(a, b, c), (a, c, d)]
    idx = np.array(idx, np.int32)
    # verify inward normals
    v0, v1, v2 = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    centre = (v0 + v1 + v2) / 3.0
    assert np.all((n * centre).sum(1) < 0), "box normals must point inward"
    return p, idx


def icosphere_geodesic(freq):
    """Geodesic sphere (class I, frequency `freq`) on the unit sphere: 20*freq^2 triangles,
    shared vertices, outward winding; returns points, indices, per-vertex normals."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    points, lookup, tris = [], {}, []

    def vid(key, pos):
        if key not in lookup:
            lookup[key] = len(points)
            points.append(pos / np.linalg.norm(pos))
        return lookup[key]

    for f in faces:
        A, B, Cc = v[f[0]], v[f[1]], v[f[2]]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                pos = (A * k + B * i + Cc * j) / freq
                # canonical key: barycentric weights on the global vertex ids (shared edges)
                w = {}
                for vert, wt in ((f[0], k), (f[1], i), (f[2], j)):
                    if wt:
                        w[vert] = w.get(vert, 0) + wt
                key = tuple(sorted(w.items()))
                grid[(i, j)] = vid(key, pos)
        for i in range(freq):
            for j in range(freq - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < freq - 1:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    p = np.array(points, np.float64)
    idx = np.array(tris, np.int32)
    v0, v1, v2 = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    flip = (n * (v0 + v1 + v2)).sum(1) < 0
    idx[flip] = idx[flip][:, [0, 2, 1]]
    assert len(idx) == 20 * freq * freq
    return p.astype(np.float32), idx, p.astype(np.float32)


def uv_sphere(segments=16, rings=11):
    """UV sphere: `rings` latitude bands x `segments` longitudes; 2*segments cap triangles plus
    2*segments*(rings-2) body triangles (16 x 11 -> 320). Unit radius, per-vertex normals."""
    pts = [[0.0, 1.0, 0.0]]
    for r in range(1, rings):
        th = math.pi * r / rings
        for s in range(segments):
            ph = 2 * math.pi * s / segments
            pts.append([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
    pts.append([0.0, -1.0, 0.0])
    p = np.array(pts, np.float64)
    south = len(pts) - 1
    tris = []

    def ring(r, s):
        return 1 + (r - 1) * segments + (s % segments)
    for s in range(segments):
        tris.append((0, ring(1, s + 1), ring(1, s)))
    for r in range(1, rings - 1):
        for s in range(segments):
            a, b = ring(r, s), ring(r, s + 1)
            c, d = ring(r + 1, s), ring(r + 1, s + 1)
            tris += [(a, b, d), (a, d, c)]
    for s in range(segments):
        tris.append((south, ring(rings - 1, s), ring(rings - 1, s + 1)))
    idx = np.array(tris, np.int32)
    v0, v1, v2 = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    flip = (n * (v0 + v1 + v2)).sum(1) < 0
    idx[flip] = idx[flip][:, [0, 2, 1]]
    assert len(idx) == 2 * segments + 2 * segments * (rings - 2)
    return p.astype(np.float32), idx, p.astype(np.float32)


def value_noise_heightfield(n, size, octaves, amplitude, rng):
    """n x n vertex grid over [-size/2, size/2]^2 (x, z); y = sum of `octaves` octaves of
    bilinear value noise. Returns points, indices (2 triangles per cell, upward winding) and
    per-vertex normals from central differences."""
    xs = np.linspace(-size / 2, size / 2, n)
    gx, gz = np.meshgrid(xs, xs, indexing="xy")
    u = (gx - xs[0]) / size
    w = (gz - xs[0]) / size
    h = np.zeros_like(gx)
    amp, freq = amplitude, 4
    for _ in range(octaves):
        lattice = rng.random((freq + 1, freq + 1))
        fx, fz = u * freq, w * freq
        ix = np.minimum(fx.astype(int), freq - 1)
        iz = np.minimum(fz.astype(int), freq - 1)
        tx, tz = fx - ix, fz - iz
        tx = tx * tx * (3 - 2 * tx)
        tz = tz * tz * (3 - 2 * tz)
        a = lattice[iz, ix] * (1 - tx) + lattice[iz, ix + 1] * tx
        b = lattice[iz + 1, ix] * (1 - tx) + lattice[iz + 1, ix + 1] * tx
        h += amp * (a * (1 - tz) + b * tz)
        amp *= 0.5
        freq *= 2
    pts = np.stack([gx, h, gz], -1).reshape(-1, 3)
    i = np.arange(n - 1)
    r0 = (i[:, None] * n + i[None, :]).reshape(-1)
    a, b, c, d = r0, r0 + 1, r0 + n, r0 + n + 1
    idx = np.empty((len(r0) * 2, 3), np.int32)
    idx[0::2] = np.stack([a, c, b], 1)
    idx[1::2] = np.stack([b, c, d], 1)
    dh_dx = np.gradient(h, xs, axis=1)
    dh_dz = np.gradient(h, xs, axis=0)
    nrm = np.stack([-dh_dx, np.ones_like(h), -dh_dz], -1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    v0, v1, v2 = pts[idx[:, 0]], pts[idx[:, 1]], pts[idx[:, 2]]
    assert np.all(np.cross(v1 - v0, v2 - v0)[:, 1] > 0)
    return pts.astype(np.float32), idx, nrm.astype(np.float32), h


# ----------------------------------------------------------------------------- cameras


def look_at_view(eye, target, up=(0.0, 1.0, 0.0)):
    """World->camera GfMatrix4d (row-vector convention, camera looks down -Z)."""
    eye, target, up = (np.asarray(a, np.float64) for a in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    upv = np.cross(back, right)
    v = np.eye(4)
    v[:3, 0], v[:3, 1], v[:3, 2] = right, upv, back
    v[3, :3] = [-right @ eye, -upv @ eye, -back @ eye]
    return v


def look_at_origin_from_sphere(rng, radius):
    """A camera at a random point of a sphere of `radius` around the origin (above the ground plane),
    looking at the origin: the multi-view workload's cameras."""
    v = rng.normal(size=3)
    v[1] = abs(v[1]) + 0.3  # above the ground
    v = v / np.linalg.norm(v) * radius
    return look_at_view(tuple(v), (0.0, 0.0, 0.0))


def perspective(fovy_deg, aspect, near=0.1, far=1000.0):
    """GfFrustum-style perspective projection, row-vector convention (clip = p * P)."""
    ys = 1.0 / math.tan(math.radians(fovy_deg) / 2)
    xs = ys / aspect
    p = np.zeros((4, 4))
    p[0, 0], p[1, 1] = xs, ys
    p[2, 2] = -(far + near) / (far - near)
    p[2, 3] = -1.0
    p[3, 2] = -2 * far * near / (far - near)
    return p


# ----------------------------------------------------------------------------- scene container


class Scene:
    """Owns the numpy arrays and exposes `.desc` (an abi.SceneDesc)."""

    def __init__(self, name, width, height):
        self.name, self.width, self.height = name, width, height
        self.meshes, self.materials, self.instances, self.lights = [], [], [], []
        self.environment = -1
        self._keep = []
        self.cameras = []  # list of (view, proj) per frame

    # -- builders
    def add_mesh(self, points, indices, normals=None, uvs=None, colors=None):
        points = np.ascontiguousarray(points, np.float32)
        indices = np.ascontiguousarray(indices, np.int32)
        self.meshes.append(dict(points=points, indices=indices,
                                normals=None if normals is None else np.ascontiguousarray(normals, np.float32),
                                uvs=None if uvs is None else np.ascontiguousarray(uvs, np.float32),
                                colors=None if colors is None else np.ascontiguousarray(colors, np.float32)))
        return len(self.meshes) - 1

    def add_material(self, **kw):
        m = dict(diffuse_color=(1, 1, 1), metallic=0.0, roughness=0.5, specular_color=(1, 1, 1),
                 specular=1.0, opacity=1.0, ior=1.5, transmission=0.0, transmission_color=(1, 1, 1),
                 transmission_depth=0.0, emission_color=(1, 1, 1), emission=0.0, coat=0.0,
                 coat_color=(1, 1, 1), coat_roughness=0.1, coat_ior=1.5, sheen=0.0, sheen_color=(1, 1, 1),
                 sheen_roughness=0.3, subsurface=0.0, subsurface_color=(1, 1, 1), thin_walled=0)
        m.update(kw)
        self.materials.append(m)
        return len(self.materials) - 1

    def add_instance(self, mesh, material, transform=None, light=-1):
        t = np.eye(4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32)
        self.instances.append(dict(mesh=mesh, material=material, transform=t, light=light))
        return len(self.instances) - 1

    def add_light(self, type_, color, intensity, transform, width=1.0, height=1.0, exposure=0.0,
                  cone_angle=180.0, cone_softness=0.0, dome_pixels=None):
        self.lights.append(dict(type=type_, color=np.asarray(color, np.float32), intensity=float(intensity),
                                exposure=float(exposure), transform=np.asarray(transform, np.float64),
                                width=float(width), height=float(height), cone_angle=float(cone_angle),
                                cone_softness=float(cone_softness),
                                dome=None if dome_pixels is None else np.ascontiguousarray(dome_pixels, np.float32)))
        return len(self.lights) - 1

    def changed(self):
        """Drops the cached descriptor after the lists were edited (instances, materials, lights);
        the next `.desc` describes the edited scene. Arrays of earlier descriptors stay alive."""
        if hasattr(self, "_desc"):
            del self._desc

    # -- stats
    def triangle_counts(self):
        unique = sum(len(m["indices"]) for m in self.meshes)
        instanced = sum(len(self.meshes[i["mesh"]]["indices"]) for i in self.instances)
        return unique, instanced

    # -- ABI view
    @property
    def desc(self):
        if hasattr(self, "_desc"):
            return self._desc
        keep = self._keep

        def fptr(a):
            if a is None:
                return C.cast(None, abi.FP), 0
            keep.append(a)
            return a.ctypes.data_as(abi.FP), a.shape[0]

        md = (abi.MeshDesc * max(1, len(self.meshes)))()
        for k, m in enumerate(self.meshes):
            d = md[k]
            d.struct_size = C.sizeof(abi.MeshDesc)
            d.points, d.num_points = fptr(m["points"])
            keep.append(m["indices"])
            d.indices, d.num_triangles = m["indices"].ctypes.data_as(abi.IP), m["indices"].shape[0]
            d.normals, d.num_normals = fptr(m["normals"])
            d.uvs, d.num_uvs = fptr(m["uvs"])
            d.colors, d.num_colors = fptr(m["colors"])
        mat = (abi.MaterialDesc * max(1, len(self.materials)))()
        for k, m in enumerate(self.materials):
            mat[k].struct_size = C.sizeof(abi.MaterialDesc)
            for name, _ in abi.MaterialDesc._fields_[1:]:
                val = m[name]
                if isinstance(val, (tuple, list, np.ndarray)):
                    arr = getattr(mat[k], name)
                    for j in range(3):
                        arr[j] = float(val[j])
                else:
                    setattr(mat[k], name, val)
        inst = (abi.InstanceDesc * max(1, len(self.instances)))()
        for k, m in enumerate(self.instances):
            inst[k].struct_size = C.sizeof(abi.InstanceDesc)
            for j in range(16):
                inst[k].transform[j] = float(m["transform"].flat[j])
            inst[k].mesh, inst[k].material, inst[k].light = m["mesh"], m["material"], m["light"]
        li = (abi.LightDesc * max(1, len(self.lights)))()
        for k, m in enumerate(self.lights):
            d = li[k]
            d.struct_size = C.sizeof(abi.LightDesc)
            d.type = m["type"]
            for j in range(3):
                d.color[j] = float(m["color"][j])
            d.intensity, d.exposure = m["intensity"], m["exposure"]
            for j in range(16):
                d.transform[j] = float(m["transform"].flat[j])
            d.width, d.height = m["width"], m["height"]
            d.shaping_cone_angle, d.shaping_cone_softness = m["cone_angle"], m["cone_softness"]
            if m["dome"] is not None:
                keep.append(m["dome"])
                d.dome_pixels = m["dome"].ctypes.data_as(abi.FP)
                d.dome_height, d.dome_width = m["dome"].shape[0], m["dome"].shape[1]
        keep += [md, mat, inst, li]
        s = abi.SceneDesc()
        s.struct_size = C.sizeof(abi.SceneDesc)
        s.meshes, s.num_meshes = C.cast(md, C.POINTER(abi.MeshDesc)), len(self.meshes)
        s.materials, s.num_materials = C.cast(mat, C.POINTER(abi.MaterialDesc)), len(self.materials)
        s.instances, s.num_instances = C.cast(inst, C.POINTER(abi.InstanceDesc)), len(self.instances)
        s.lights, s.num_lights = C.cast(li, C.POINTER(abi.LightDesc)), len(self.lights)
        s.environment = self.environment
        self._desc = s
        return s

    def frame(self, frame_index, camera_index=None):
        view, proj = self.cameras[frame_index if camera_index is None else camera_index]
        return abi.frame(view, proj, frame_index)


def _translate_scale(t, s):
    m = np.eye(4)
    m[0, 0] = m[1, 1] = m[2, 2] = s
    m[3, :3] = t
    return m


def _random_material(scene, rng):
    return scene.add_material(diffuse_color=tuple(rng.uniform(0.1, 0.9, 3)),
                              roughness=float(rng.uniform(0.1, 0.9)),
                              metallic=float(rng.integers(0, 2)))


def _rect_light_down(pos, yaw):
    """Rect light whose -Z axis points straight down (-Y), rotated by `yaw` about Y."""
    cy, sy = math.cos(yaw), math.sin(yaw)
    m = np.eye(4)
    m[0, :3] = [cy, 0.0, -sy]      # local X
    m[1, :3] = [-sy, 0.0, -cy]     # local Y
    m[2, :3] = [0.0, 1.0, 0.0]     # local Z -> +Y, so local -Z faces down
    m[3, :3] = pos
    return m


def box_spheres(width=256, height=256, seed=0, frames=16):
    """Closed box + two icospheres + one rect light (the CPU-runnable path-tracer/RIS config)."""
    rng = np.random.default_rng(seed)
    sc = Scene("box_spheres", width, height)
    bp, bi = box_mesh(1.0)
    m_box = sc.add_mesh(bp, bi)
    sp, si, sn = icosphere_geodesic(11)
    m_sph = sc.add_mesh(sp, si, normals=sn)
    sc.add_instance(m_box, _random_material(sc, rng))
    for _ in range(2):
        r = float(rng.uniform(0.25, 0.4))
        c = (float(rng.uniform(-0.5, 0.5)), -1.0 + r, float(rng.uniform(-0.6, 0.2)))
        sc.add_instance(m_sph, _random_material(sc, rng), _translate_scale(c, r))
    sc.add_light(abi.HDR_LIGHT_RECT, blackbody_rgb(6500.0), 20.0, _rect_light_down((0.0, 0.98, 0.0), 0.0),
                 width=0.5, height=0.5)
    view = look_at_view((0.0, 0.0, 0.95), (0.0, -0.2, -1.0))
    proj = perspective(60.0, width / height, 0.01, 100.0)
    sc.cameras = [(view, proj)] * frames
    return sc


def many_lights(width=1920, height=1080, seed=0, frames=64, n_lights=4096, n_spheres=500, grid=512,
                dolly=True, sphere_rings=11, intensity_decades=4.0, dome=None, name="many_lights"):
    """Height field + instanced spheres + many point/rect lights (the 1080p ReSTIR config).
    dome=(w, h, sigma) adds an environment map with log-normal texels exp(N(0, sigma))."""
    rng = np.random.default_rng(seed)
    sc = Scene(name, width, height)
    size = 100.0
    hp, hi, hn, hgt = value_noise_heightfield(grid, size, 4, 6.0, rng)
    m_h = sc.add_mesh(hp, hi, normals=hn)
    sc.add_instance(m_h, _random_material(sc, rng))
    up, ui, un = uv_sphere(16, sphere_rings)
    m_s = sc.add_mesh(up, ui, normals=un)
    xs = np.linspace(-size / 2, size / 2, grid)

    def height_at(x, z):
        ix = int(np.clip(np.searchsorted(xs, x), 0, grid - 1))
        iz = int(np.clip(np.searchsorted(xs, z), 0, grid - 1))
        return float(hgt[iz, ix])
    for _ in range(n_spheres):
        r = float(rng.uniform(0.5, 2.0))
        x, z = (float(v) for v in rng.uniform(-size * 0.45, size * 0.45, 2))
        sc.add_instance(m_s, _random_material(sc, rng), _translate_scale((x, height_at(x, z) + r * 0.8, z), r))
    n_point = n_lights // 2
    for k in range(n_lights):
        x, z = (float(v) for v in rng.uniform(-size * 0.48, size * 0.48, 2))
        y = height_at(x, z) + float(rng.uniform(2.0, 10.0))
        inten = float(10.0 ** rng.uniform(0.0, intensity_decades))
        col = blackbody_rgb(float(rng.uniform(2000.0, 10000.0)))
        if k < n_point:
            sc.add_light(abi.HDR_LIGHT_POINT, col, inten, _translate_scale((x, y, z), 1.0))
        else:
            side = float(rng.uniform(0.05, 0.5))
            sc.add_light(abi.HDR_LIGHT_RECT, col, inten, _rect_light_down((x, y, z), float(rng.uniform(0, math.pi))),
                         width=side, height=side)
    if dome is not None:
        dw, dh, sigma = dome
        tex = np.exp(rng.normal(0.0, sigma, (dh, dw, 3))).astype(np.float32)
        sc.environment = sc.add_light(abi.HDR_LIGHT_DOME, (1.0, 1.0, 1.0), 1.0, np.eye(4), dome_pixels=tex)
    proj = perspective(55.0, width / height, 0.1, 1000.0)
    cams = []
    for f in range(frames):
        t = f / max(1, frames - 1) if dolly else 0.0
        eye = np.array([-30.0 + 60.0 * t, 22.0, 45.0])
        cams.append((look_at_view(eye, eye + np.array([0.0, -14.0, -40.0])), proj))
    sc.cameras = cams
    return sc


def large_scene(width=1920, height=1080, seed=0, frames=8, n_lights=65536, n_meshes=4000, grid=1024,
                dolly=False, dome=(2048, 1024, 2.0)):
    """The large-scene workload: a 1024x1024-vertex displaced grid (2,093,058 triangles) plus 4000
    instances of a 256-triangle sphere, 65,536 point / rect lights with intensities over 5 decades,
    and a 2048x1024 environment map with log-normal texels exp(N(0, 2))."""
    return many_lights(width, height, seed, frames, n_lights, n_meshes, grid, dolly, sphere_rings=9,
                       intensity_decades=5.0, dome=dome, name="large")


CONFIGS = {"box_spheres": box_spheres, "many_lights": many_lights, "large": large_scene}


def _dome_texels(w, h, rng):
    """Small environment map: log-normal texels (mu 0, sigma 1) times a sky gradient."""
    tex = np.exp(rng.normal(0.0, 1.0, (h, w, 3))).astype(np.float32)
    grad = np.linspace(1.5, 0.2, h, dtype=np.float32)[:, None, None]
    return (tex * grad * np.array([0.6, 0.7, 1.0], np.float32)).astype(np.float32)


def features(width=48, height=48, seed=3, frames=4, dome=True, dof=False, emissive=True, opaque=False,
             mesh_light=None):
    """Every shading feature of the path in one small scene: coat / sheen / transmission with Beer
    absorption / partial opacity / emission / DefaultMaterial with vertex colours; rect, cone-shaped
    point, distant and dome lights (the dome is the environment); an emissive sphere registered as a
    mesh light. dome / emissive / opaque=True select the variant whose BSDF rays reach no emitter
    and whose materials are all opaque (the deferred-shadow path at depth 1 with skipVisibility).
    mesh_light=False keeps the emissive material but registers no mesh light (default: emissive)."""
    rng = np.random.default_rng(seed)
    sc = Scene("features", width, height)
    bp, bi = box_mesh(2.0)
    # open the box's +z wall (camera side) so the dome is visible
    keep = np.ones(len(bi), bool)
    keep[2:4] = False
    m_box = sc.add_mesh(bp, bi[keep])
    sp, si, sn = icosphere_geodesic(3)
    colors = np.abs(sp).astype(np.float32)
    m_sph = sc.add_mesh(sp, si, normals=sn, colors=colors)
    sc.add_instance(m_box, sc.add_material(diffuse_color=(0.6, 0.6, 0.55), roughness=0.7))
    mats = [
        dict(diffuse_color=(0.8, 0.2, 0.2), roughness=0.4, coat=0.8, coat_roughness=0.2, sheen=0.5),
        dict(diffuse_color=(0.9, 0.9, 0.9), roughness=0.3, transmission=0.9, transmission_depth=0.5,
             transmission_color=(0.5, 0.8, 0.9), ior=1.45),
        dict(diffuse_color=(0.2, 0.7, 0.3), roughness=0.6, opacity=1.0 if opaque else 0.5),
        dict(diffuse_color=(0.9, 0.8, 0.3), roughness=0.2, metallic=1.0),
        dict(diffuse_color=(0.1, 0.1, 0.1), roughness=0.9, emission=4.0 if emissive else 0.0,
             emission_color=(1.0, 0.6, 0.3)),
        None,  # DefaultMaterial: vertex colours, Lambert sampling
    ]
    pos = [(-1.1, -1.4, -0.8), (0.0, -1.4, -1.0), (1.1, -1.4, -0.8), (-0.6, -1.5, 0.4), (0.7, -1.5, 0.5),
           (0.0, -1.55, 1.1)]
    light_idx = {}
    for k, (m, p) in enumerate(zip(mats, pos)):
        mat = sc.add_material(**m) if m is not None else -1
        inst = sc.add_instance(m_sph, mat, _translate_scale(p, 0.45))
        if k == 4:
            light_idx["emissive"] = inst
    rect = sc.add_light(abi.HDR_LIGHT_RECT, blackbody_rgb(5000.0), 15.0, _rect_light_down((0.3, 1.95, -0.2), 0.4),
                        width=0.6, height=0.4)
    cone = _translate_scale((-1.2, 1.6, 0.6), 1.0)
    cone[:3, :3] = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)  # local -Z -> world -Y
    sc.add_light(abi.HDR_LIGHT_POINT, (1.0, 0.9, 0.7), 2.0, cone, cone_angle=50.0, cone_softness=0.3)
    dist = np.eye(4)
    d = np.array([0.3, -0.8, -0.5])
    d /= np.linalg.norm(d)
    z = -d
    x = np.cross([0, 1, 0], z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    dist[0, :3], dist[1, :3], dist[2, :3] = x, y, z
    sc.add_light(abi.HDR_LIGHT_DISTANT, (1.0, 1.0, 0.9), 0.8, dist, exposure=-1.0)
    # the emissive sphere is the mesh prim of this (small) rect light: GetLightAtHit finds it
    em = sc.add_light(abi.HDR_LIGHT_RECT, (1.0, 0.6, 0.3), 4.0, _rect_light_down(pos[4], 0.0), width=0.3, height=0.3)
    if emissive if mesh_light is None else mesh_light:
        sc.instances[light_idx["emissive"]]["light"] = em
    if dome:
        sc.environment = sc.add_light(abi.HDR_LIGHT_DOME, (1.0, 1.0, 1.0), 0.7, np.eye(4),
                                      dome_pixels=_dome_texels(32, 16, rng))
    view = look_at_view((0.2, -0.2, 3.6), (0.0, -0.9, -0.5))
    proj = perspective(50.0, width / height, 0.05, 100.0)
    sc.cameras = [(view, proj)] * frames
    sc.camera_params = None
    if dof:
        cp = abi.CameraParams()
        cp.enable_dof, cp.enable_exposure = 1, 1
        cp.focal_length, cp.f_stop, cp.focus_distance, cp.bokeh_blades = 35.0, 8.0, 4.0, 6
        cp.lens_distortion, cp.iso, cp.shutter_speed = 0.05, 200.0, 0.04
        sc.camera_params = cp
    return sc


def frame_with_camera(sc, frame_index):
    f = sc.frame(frame_index)
    if getattr(sc, "camera_params", None) is not None:
        f.has_camera_params = 1
        f.camera = sc.camera_params
    return f


CONFIGS["features"] = features
