"""Scene description and the synthetic benchmark scenes.

A scene is world-space triangle geometry (mesh instances are flattened when added), a material
table of 20-float records, optional RGB fp32 textures and a pinhole camera. This mirrors the
parts of Falcor's SceneBuilder/TriangleMesh/Transform/Camera that the two passes read
(Source/Falcor/Scene/SceneBuilder.cpp, TriangleMesh.cpp:58-195, Transform.cpp:80-100,
Camera.cpp:112-193).

Material record layout (float32[20]):
  0 type (0 StandardMaterial, 1 PBRTDiffuseMaterial)   1-3 baseColor   4 roughness   5 metallic
  6 ior   7-9 emissive   10 emissiveFactor   11 doubleSided   12 baseColor texture (-1 none)
  13 emissive texture + 1 (0 none)   14 specularTransmission   15 diffuseTransmission
  16-18 transmission colour   19 thinSurface
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MAT_STANDARD = 0
MAT_PBRT_DIFFUSE = 1
MAT_STRIDE = 20


# --------------------------------------------------------------------------- meshes
@dataclass
class TriangleMesh:
    """Object-space mesh (TriangleMesh.cpp). Positions/normals float32 [V,3], uv [V,2], indices [T,3]."""

    positions: np.ndarray
    normals: np.ndarray
    texcoords: np.ndarray
    indices: np.ndarray

    @staticmethod
    def create_quad(size=(1.0, 1.0)) -> "TriangleMesh":
        # TriangleMesh.cpp:58-78
        hx, hy = 0.5 * size[0], 0.5 * size[1]
        p = np.array([[-hx, 0, -hy], [hx, 0, -hy], [-hx, 0, hy], [hx, 0, hy]], np.float32)
        n = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
        uv = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32)
        idx = np.array([[2, 1, 0], [1, 2, 3]], np.uint32)
        return TriangleMesh(p, n, uv, idx)

    @staticmethod
    def create_cube(size=(1.0, 1.0, 1.0)) -> "TriangleMesh":
        # TriangleMesh.cpp:102-149
        P = [
            [[-0.5, -0.5, -0.5], [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, -0.5, -0.5]],
            [[-0.5, 0.5, 0.5], [-0.5, 0.5, -0.5], [0.5, 0.5, -0.5], [0.5, 0.5, 0.5]],
            [[-0.5, 0.5, -0.5], [-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5]],
            [[0.5, 0.5, 0.5], [0.5, -0.5, 0.5], [-0.5, -0.5, 0.5], [-0.5, 0.5, 0.5]],
            [[-0.5, 0.5, 0.5], [-0.5, -0.5, 0.5], [-0.5, -0.5, -0.5], [-0.5, 0.5, -0.5]],
            [[0.5, 0.5, -0.5], [0.5, -0.5, -0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5]],
        ]
        N = [[0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0]]
        UV = [[0, 0], [1, 0], [1, 1], [0, 1]]
        s = np.asarray(size, np.float32)
        sign = np.where(s < 0, -1.0, 1.0).astype(np.float32)
        pos, nrm, uv, idx = [], [], [], []
        for i in range(6):
            b = len(pos)
            idx += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
            for j in range(4):
                pos.append(np.asarray(P[i][j], np.float32) * s)
                nrm.append(np.asarray(N[i], np.float32) * sign)
                uv.append(UV[j])
        return TriangleMesh(np.array(pos, np.float32), np.array(nrm, np.float32), np.array(uv, np.float32),
                            np.array(idx, np.uint32))

    @staticmethod
    def create_sphere(radius=0.5, segments_u=32, segments_v=16) -> "TriangleMesh":
        # TriangleMesh.cpp:151-191 (UV sphere; createSphere(r, 100, 100) has 10,201 vertices, 20,000 triangles)
        pos, nrm, uv = [], [], []
        for v in range(segments_v + 1):
            for u in range(segments_u + 1):
                uvx, uvy = np.float32(u / float(segments_u)), np.float32(v / float(segments_v))
                theta = uvx * np.float32(2.0) * np.float32(math.pi)
                phi = uvy * np.float32(math.pi)
                d = np.array([math.cos(theta) * math.sin(phi), math.cos(phi), math.sin(theta) * math.sin(phi)], np.float32)
                pos.append(d * np.float32(radius))
                nrm.append(d)
                uv.append([uvx, uvy])
        idx = []
        for v in range(segments_v):
            for u in range(segments_u):
                i0 = v * (segments_u + 1) + u
                i1 = v * (segments_u + 1) + (u + 1) % (segments_u + 1)
                i2 = (v + 1) * (segments_u + 1) + u
                i3 = (v + 1) * (segments_u + 1) + (u + 1) % (segments_u + 1)
                idx += [[i0, i1, i2], [i2, i1, i3]]
        return TriangleMesh(np.array(pos, np.float32), np.array(nrm, np.float32), np.array(uv, np.float32),
                            np.array(idx, np.uint32))

    @staticmethod
    def create_icosphere(radius=1.0, subdivisions=3) -> "TriangleMesh":
        """Icosahedron with `subdivisions` midpoint refinements (3 -> 1280 triangles, 642 vertices).

        Spherical texture coordinates u = atan2(z, x) / 2pi + 0.5, v = acos(y) / pi.
        """
        t = (1.0 + 5.0 ** 0.5) / 2.0
        verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                 (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
        verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
        faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                 (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
                 (6, 2, 10), (8, 6, 7), (9, 8, 1)]
        for _ in range(subdivisions):
            cache = {}

            def mid(a, b):
                key = (min(a, b), max(a, b))
                if key not in cache:
                    m = verts[a] + verts[b]
                    verts.append(m / np.linalg.norm(m))
                    cache[key] = len(verts) - 1
                return cache[key]

            nf = []
            for a, b, c in faces:
                ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
                nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
            faces = nf
        d = np.array(verts, np.float64)
        # Orient faces outward.
        f = np.array(faces, np.int64)
        cr = np.cross(d[f[:, 1]] - d[f[:, 0]], d[f[:, 2]] - d[f[:, 0]])
        flip = (cr * d[f].mean(axis=1)).sum(axis=1) < 0
        f[flip] = f[flip][:, [0, 2, 1]]
        u = np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi) + 0.5
        v = np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi
        return TriangleMesh((d * radius).astype(np.float32), d.astype(np.float32),
                            np.stack([u, v], 1).astype(np.float32), f.astype(np.uint32))


# --------------------------------------------------------------------------- transforms
def _rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


@dataclass
class Transform:
    """Scale-rotate-translate transform (Transform.cpp:80-100, default composition T*R*S)."""

    scaling: tuple = (1.0, 1.0, 1.0)
    translation: tuple = (0.0, 0.0, 0.0)
    rotation_euler: tuple = (0.0, 0.0, 0.0)  # radians, applied X then Y then Z
    rotation_matrix: np.ndarray | None = None

    def matrices(self):
        s = np.asarray(self.scaling if np.ndim(self.scaling) else (self.scaling,) * 3, np.float64)
        if self.rotation_matrix is not None:
            r = np.asarray(self.rotation_matrix, np.float64)
        else:
            ax, ay, az = self.rotation_euler
            r = _rot_z(az) @ _rot_y(ay) @ _rot_x(ax)
        return r, s, np.asarray(self.translation, np.float64)

    def matrix(self) -> np.ndarray:
        """4x4 local-to-world matrix T * R * S (float64)."""
        r, s, t = self.matrices()
        m = np.eye(4)
        m[:3, :3] = r * s[None, :]
        m[:3, 3] = t
        return m

    @staticmethod
    def deg(scaling=(1, 1, 1), translation=(0, 0, 0), rotation_euler_deg=(0, 0, 0)) -> "Transform":
        return Transform(scaling, translation, tuple(math.radians(a) for a in rotation_euler_deg))


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform (Haar) random rotation from a normalized Gaussian quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# --------------------------------------------------------------------------- materials
# Material record fields (20 floats; see include/fhp.h).
M_TYPE, M_BASE, M_ROUGH, M_METAL, M_IOR, M_EMIT, M_EMIT_FACTOR, M_DOUBLE_SIDED, M_TEX = 0, 1, 4, 5, 6, 7, 10, 11, 12
M_EMIT_TEX = 13  # emissive texture index + 1 (0: the emissive colour)
M_SPEC_TRANS, M_DIFF_TRANS, M_TRANS, M_THIN = 14, 15, 16, 19


def standard_material(base=(1, 1, 1), roughness=0.0, metallic=0.0, ior=1.5, emissive=(0, 0, 0), emissive_factor=1.0,
                      double_sided=False, texture=-1, emissive_texture=-1, specular_transmission=0.0,
                      diffuse_transmission=0.0, transmission=(1.0, 1.0, 1.0), thin_surface=False) -> np.ndarray:
    """StandardMaterial record (defaults of BasicMaterialData.slang:52-57: baseColor 1, specular 0)."""
    m = np.zeros(MAT_STRIDE, np.float32)
    m[0] = MAT_STANDARD
    m[1:4] = base
    m[4] = roughness
    m[5] = metallic
    m[6] = ior
    m[7:10] = emissive
    m[10] = emissive_factor
    m[11] = 1.0 if double_sided else 0.0
    m[12] = texture
    m[M_EMIT_TEX] = emissive_texture + 1
    m[M_SPEC_TRANS] = specular_transmission
    m[M_DIFF_TRANS] = diffuse_transmission
    m[M_TRANS:M_TRANS + 3] = transmission
    m[M_THIN] = 1.0 if thin_surface else 0.0
    return m


def pbrt_diffuse_material(albedo=(0.5, 0.5, 0.5), double_sided=False, texture=-1) -> np.ndarray:
    """PBRTDiffuseMaterial record (Lambertian; PBRTDiffuseMaterialInstance.slang:40-61)."""
    m = np.zeros(MAT_STRIDE, np.float32)
    m[0] = MAT_PBRT_DIFFUSE
    m[1:4] = albedo
    m[6] = 1.5
    m[10] = 1.0
    m[11] = 1.0 if double_sided else 0.0
    m[12] = texture
    m[M_TRANS:M_TRANS + 3] = 1.0
    return m


# --------------------------------------------------------------------------- analytic lights
LIGHT_POINT, LIGHT_DIRECTIONAL = 0, 1


def point_light(intensity=(1.0, 1.0, 1.0), position=(0.0, 0.0, 0.0), direction=(0.0, -1.0, 0.0),
                opening_angle=math.pi, penumbra_angle=0.0) -> np.ndarray:
    """PointLight record (LightData.slang:48-66 defaults; a spot light when opening_angle < pi)."""
    r = np.zeros(16, np.float32)
    r[0] = LIGHT_POINT
    r[1:4] = intensity
    r[4:7] = position
    r[7] = opening_angle
    r[8:11] = direction
    r[11] = penumbra_angle
    return r


def directional_light(intensity=(1.0, 1.0, 1.0), direction=(0.0, -1.0, 0.0)) -> np.ndarray:
    """DirectionalLight record: irradiance `intensity` arriving along `direction`."""
    r = np.zeros(16, np.float32)
    r[0] = LIGHT_DIRECTIONAL
    r[1:4] = intensity
    r[8:11] = direction
    return r


# --------------------------------------------------------------------------- scene data
@dataclass
class Camera:
    position: tuple = (0.0, 0.0, 0.0)
    target: tuple = (0.0, 0.0, -1.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov_y: float = math.radians(40.0)
    aspect: float = 1.0
    focal_distance: float = 10000.0  # CameraData.slang:63
    near_z: float = 0.1               # CameraData.slang:53
    far_z: float = 1000.0             # CameraData.slang:55

    def pack(self) -> np.ndarray:
        return np.array([*self.position, *self.target, *self.up, self.fov_y, self.aspect, self.focal_distance,
                         self.near_z, self.far_z], np.float32)


@dataclass
class SceneData:
    positions: np.ndarray
    normals: np.ndarray
    texcoords: np.ndarray
    indices: np.ndarray
    tri_material: np.ndarray
    tri_mesh: np.ndarray
    materials: np.ndarray
    camera: Camera
    tex_desc: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    texels: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    mesh_names: list = field(default_factory=list)
    mesh_vertex_offsets: np.ndarray | None = None  # mesh m owns vertices [off[m], off[m + 1])
    # Environment map: {"texels": [H, W, 3], "intensity", "tint", "rotation" (degrees XYZ)} or None.
    env_map: dict | None = None
    # Local-to-world matrix (4x4) each mesh was placed with (identity when built from world data).
    mesh_transforms: list = field(default_factory=list)
    # Analytic lights: records of point_light() / directional_light() (16 floats each).
    analytic_lights: list = field(default_factory=list)

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])

    def camera_array(self) -> np.ndarray:
        return self.camera.pack()


class SceneBuilder:
    """Flattens mesh instances into world-space arrays (SceneBuilder::addMeshInstance semantics)."""

    def __init__(self):
        self._pos, self._nrm, self._uv, self._idx, self._mat, self._mesh = [], [], [], [], [], []
        self.materials: list[np.ndarray] = []
        self.textures: list[np.ndarray] = []
        self.mesh_names: list[str] = []
        self._nv = 0
        self._mesh_off = [0]
        self._xf: list[np.ndarray] = []
        self.lights: list[np.ndarray] = []
        self.camera = Camera()

    def add_material(self, rec: np.ndarray) -> int:
        self.materials.append(np.asarray(rec, np.float32))
        return len(self.materials) - 1

    def add_light(self, rec: np.ndarray) -> int:
        """Adds an analytic light record (point_light / directional_light)."""
        self.lights.append(np.asarray(rec, np.float32).copy())
        return len(self.lights) - 1

    def add_texture(self, rgb: np.ndarray) -> int:
        rgb = np.ascontiguousarray(rgb, np.float32)
        assert rgb.ndim == 3 and rgb.shape[2] == 3
        self.textures.append(rgb)
        return len(self.textures) - 1

    def add_mesh_instance(self, mesh: TriangleMesh, material, transform: Transform | None = None, name: str = "") -> int:
        """Adds a mesh instance; `material` is one id or an array of per-triangle ids. Returns the mesh id."""
        t = transform or Transform()
        r, s, tr = t.matrices()
        p = (mesh.positions.astype(np.float64) * s) @ r.T + tr
        nm = (mesh.normals.astype(np.float64) / s) @ r.T
        nm /= np.maximum(np.linalg.norm(nm, axis=1, keepdims=True), 1e-30)
        idx = mesh.indices.astype(np.int64)
        if np.prod(s) < 0:  # negative determinant flips the winding
            idx = idx[:, [0, 2, 1]]
        mesh_id = len(self.mesh_names)
        self.mesh_names.append(name or f"mesh{mesh_id}")
        self._pos.append(p.astype(np.float32))
        self._nrm.append(nm.astype(np.float32))
        self._uv.append(mesh.texcoords.astype(np.float32))
        self._idx.append((idx + self._nv).astype(np.uint32))
        mat = np.broadcast_to(np.asarray(material, np.uint32), (idx.shape[0],))
        self._mat.append(mat.copy())
        self._mesh.append(np.full(idx.shape[0], mesh_id, np.uint32))
        self._xf.append(t.matrix())
        self._nv += p.shape[0]
        self._mesh_off.append(self._nv)
        return mesh_id

    def build(self) -> SceneData:
        tex_desc = np.zeros((len(self.textures), 4), np.int32)
        off = 0
        for i, t in enumerate(self.textures):
            tex_desc[i] = (off, t.shape[1], t.shape[0], 0)
            off += t.size
        texels = np.concatenate([t.reshape(-1) for t in self.textures]) if self.textures else np.zeros(0, np.float32)
        return SceneData(
            positions=np.concatenate(self._pos), normals=np.concatenate(self._nrm), texcoords=np.concatenate(self._uv),
            indices=np.concatenate(self._idx), tri_material=np.concatenate(self._mat),
            tri_mesh=np.concatenate(self._mesh), materials=np.stack(self.materials), camera=self.camera,
            tex_desc=tex_desc, texels=texels.astype(np.float32), mesh_names=list(self.mesh_names),
            mesh_vertex_offsets=np.asarray(self._mesh_off, np.uint32), mesh_transforms=[m.copy() for m in self._xf],
            analytic_lights=[r.copy() for r in self.lights])


# --------------------------------------------------------------------------- benchmark scenes
def closed_box(seed: int = 0, aspect: float = 1.0) -> SceneData:
    """Closed box: 5 quads + 2 rotated boxes (Lambertian albedo U[0.1, 0.8] per object) and an
    emissive ceiling quad of radiance 17 (36 triangles). Layout of Scene/Box/CornellBox.pyscene:60-123."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    quad, cube = TriangleMesh.create_quad(), TriangleMesh.create_cube()
    walls = [
        ("Floor", Transform.deg((0.55, 1.0, 0.56))),
        ("Ceiling", Transform.deg((0.55, 1.0, 0.56), (0, 0.55, 0), (180, 0, 0))),
        ("Back Wall", Transform.deg((0.55, 1.0, 0.56), (0, 0.275, -0.275), (90, 0, 0))),
        ("Left Wall", Transform.deg((0.55, 1.0, 0.56), (-0.275, 0.275, 0), (0, 0, -90))),
        ("Right Wall", Transform.deg((0.55, 1.0, 0.56), (0.275, 0.275, 0), (0, 0, 90))),
    ]
    for name, t in walls:
        m = sb.add_material(pbrt_diffuse_material(rng.uniform(0.1, 0.8, 3)))
        sb.add_mesh_instance(quad, m, t, name)
    boxes = [
        ("Large Box", Transform((0.165, 0.33, 0.165), (-0.093, 0.165, -0.071), (0, -1.27, 0))),
        ("Small Box", Transform((0.165, 0.165, 0.165), (0.09, 0.0825, 0.111), (0, -0.29, 0))),
    ]
    for name, t in boxes:
        m = sb.add_material(pbrt_diffuse_material(rng.uniform(0.1, 0.8, 3)))
        sb.add_mesh_instance(cube, m, t, name)
    light = sb.add_material(standard_material(base=(0, 0, 0), roughness=0.5, emissive=(17, 17, 17)))
    sb.add_mesh_instance(quad, light, Transform.deg((0.13, 0.13, 0.13), (0, 0.549, 0), (180, 0, 0)), "Light")
    sb.camera = Camera((0, 0.28, 1.2), (0, 0.28, 0), (0, 1, 0), math.radians(40.0), aspect)
    return sb.build()


def glass_slab(colour=(1.0, 1.0, 1.0), roughness: float = 0.0, ior: float = 1.5, thin: bool = False,
               slab: bool = True, radiance: float = 1.0, fov_deg: float = 2.0, aspect: float = 1.0) -> SceneData:
    """A 2 x 2 x 0.2 slab of a specular-transmission StandardMaterial (declared single-sided) at the
    origin, in front of a 4 x 4 emissive quad of `radiance` at z = -1 that faces the camera at
    (0, 0, 5) with a narrow field of view. At normal incidence a smooth slab transmits
    radiance * (1 - F) / (1 + F) * transmission albedo^2 (F the normal-incidence Fresnel term)."""
    sb = SceneBuilder()
    if slab:
        g = sb.add_material(standard_material(base=(1, 1, 1), roughness=roughness, ior=ior, specular_transmission=1.0,
                                              transmission=colour, thin_surface=thin))
        sb.add_mesh_instance(TriangleMesh.create_cube((2.0, 2.0, 0.2)), g, None, "Slab")
    e = sb.add_material(standard_material(base=(0, 0, 0), roughness=1.0, ior=1.0, emissive=(radiance,) * 3))
    sb.add_mesh_instance(TriangleMesh.create_quad(), e, Transform.deg((4, 1, 4), (0, 0, -1), (90, 0, 0)), "Light")
    sb.camera = Camera((0, 0, 5), (0, 0, 0), (0, 1, 0), math.radians(fov_deg), aspect)
    return sb.build()


def vertex_normals(positions: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (degenerate triangles contribute nothing)."""
    p = np.asarray(positions, np.float64)
    idx = np.asarray(indices, np.int64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)


def displaced_sphere_in_box(seed: int = 0, radius: float = 0.12, segments=(100, 100), sigma: float = 0.05,
                            center=(0.0, 0.2, 0.0), aspect: float = 1.0, front_wall: bool = False) -> SceneData:
    """A Lambertian UV sphere (createSphere(r, 100, 100): 10,201 vertices, 20,000 triangles) inside
    the closed box's walls and ceiling light. Vertices are displaced along the normal by a smooth
    seeded field with standard deviation sigma * radius (sigma = 0 gives the plain sphere); vertex
    normals are recomputed from the displaced surface. The sphere is the mesh named "Sphere".
    front_wall adds a sixth wall closing the box's open side, so that views from inside the box
    (orbit_cameras) see walls in every direction and no ray leaves the scene."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    quad = TriangleMesh.create_quad()
    walls = [
        ("Floor", Transform.deg((0.55, 1.0, 0.56))),
        ("Ceiling", Transform.deg((0.55, 1.0, 0.56), (0, 0.55, 0), (180, 0, 0))),
        ("Back Wall", Transform.deg((0.55, 1.0, 0.56), (0, 0.275, -0.275), (90, 0, 0))),
        ("Left Wall", Transform.deg((0.55, 1.0, 0.56), (-0.275, 0.275, 0), (0, 0, -90))),
        ("Right Wall", Transform.deg((0.55, 1.0, 0.56), (0.275, 0.275, 0), (0, 0, 90))),
    ]
    for name, t in walls:
        m = sb.add_material(pbrt_diffuse_material(rng.uniform(0.1, 0.8, 3)))
        sb.add_mesh_instance(quad, m, t, name)
    if front_wall:
        sb.add_mesh_instance(quad, sb.add_material(pbrt_diffuse_material((0.5, 0.5, 0.5))),
                             Transform.deg((0.55, 1.0, 0.56), (0, 0.275, 0.275), (-90, 0, 0)), "Front Wall")
    light = sb.add_material(standard_material(base=(0, 0, 0), roughness=0.5, emissive=(17, 17, 17)))
    sb.add_mesh_instance(quad, light, Transform.deg((0.13, 0.13, 0.13), (0, 0.549, 0), (180, 0, 0)), "Light")
    sph = TriangleMesh.create_sphere(1.0, segments[0], segments[1])
    n0 = sph.positions.astype(np.float64)
    d = np.zeros(n0.shape[0])
    if sigma > 0:
        centers = rng.normal(size=(8, 3))
        centers /= np.linalg.norm(centers, axis=1, keepdims=True)
        amp = rng.normal(size=8)
        for c, a in zip(centers, amp):
            d += a * np.exp(6.0 * (n0 @ c - 1.0))
        d = (d - d.mean()) / max(d.std(), 1e-12) * sigma * radius
    pos = (n0 * (radius + d)[:, None] + np.asarray(center)).astype(np.float32)
    mesh = TriangleMesh(pos, vertex_normals(pos, sph.indices), sph.texcoords, sph.indices)
    m = sb.add_material(pbrt_diffuse_material(rng.uniform(0.3, 0.8, 3)))
    sb.add_mesh_instance(mesh, m, None, "Sphere")
    sb.camera = Camera((0, 0.28, 1.2), (0, 0.28, 0), (0, 1, 0), math.radians(40.0), aspect)
    return sb.build()


def orbit_cameras(n: int = 32, center=(0.0, 0.2, 0.0), radius: float = 0.22, heights=(-0.08, 0.08),
                  fov_y: float = math.radians(70.0), aspect: float = 1.0) -> list:
    """n pinhole cameras on rings around `center` (len(heights) rings of n / len(heights) cameras
    at the given heights above it), each looking at the centre: the views of the batched
    vertex-gradient workload around the sphere of displaced_sphere_in_box, inside the box."""
    per = n // len(heights)
    cams = []
    for r, dy in enumerate(heights):
        for k in range(per):
            a = 2.0 * math.pi * (k + 0.5 * r) / per
            pos = (center[0] + radius * math.sin(a), center[1] + dy, center[2] + radius * math.cos(a))
            cams.append(Camera(pos, tuple(center), (0.0, 1.0, 0.0), fov_y, aspect))
    return cams


def sphere_field(seed: int = 1, num_spheres: int = 1024, num_emissive: int = 4096, box: float = 100.0,
                 subdivisions: int = 3, aspect: float = 16 / 9, camera: Camera | None = None) -> SceneData:
    """Icosphere field: log-normal(0, 0.5) scales, Haar rotations, uniform positions in [0, box]^3,
    StandardMaterial (baseColor U[0.1,0.9], roughness U[0.05,1], metallic Bernoulli(0.3)) and
    `num_emissive` triangles with log-normal(0, 1) white radiance."""
    rng = np.random.default_rng(seed)
    ico = TriangleMesh.create_icosphere(1.0, subdivisions)
    sb = SceneBuilder()
    mats = []
    for i in range(num_spheres):
        m = standard_material(base=rng.uniform(0.1, 0.9, 3), roughness=rng.uniform(0.05, 1.0),
                              metallic=float(rng.random() < 0.3))
        mats.append(sb.add_material(m))
    scales = rng.lognormal(0.0, 0.5, num_spheres)
    for i in range(num_spheres):
        t = Transform((scales[i],) * 3, tuple(rng.uniform(0, box, 3)), rotation_matrix=random_rotation(rng))
        sb.add_mesh_instance(ico, mats[i], t, f"sphere{i}")
    data = sb.build()
    ntri = data.num_triangles
    em = rng.choice(ntri, size=min(num_emissive, ntri), replace=False)
    radiance = rng.lognormal(0.0, 1.0, em.size)
    new_mats = [data.materials]
    tri_mat = data.tri_material.copy()
    base = data.materials.shape[0]
    extra = np.repeat(data.materials[tri_mat[em]], 1, axis=0).copy()
    extra[:, 7:10] = radiance[:, None]
    extra[:, 10] = 1.0
    new_mats.append(extra)
    tri_mat[em] = base + np.arange(em.size, dtype=np.uint32)
    data.materials = np.concatenate(new_mats).astype(np.float32)
    data.tri_material = tri_mat
    data.camera = camera or Camera((box * 0.5, box * 0.5, -0.4 * box), (box * 0.5, box * 0.5, box * 0.5), (0, 1, 0),
                                   math.radians(70.0), aspect)
    return data


def batch_workload(seed: int = 7, num_scenes: int = 16, num_views: int = 64, res: int = 512,
                   tri_range=(1e4, 2e6), spp_choices=(1, 2, 4, 8, 16)):
    """Batched-forward workload: `num_scenes` sphere fields whose triangle counts are log-uniform in
    `tri_range` (1280-triangle icospheres at the sphere-field density of 1024 spheres per 100^3,
    at most 4096 emissive triangles), and `num_views` views, each with a scene, a camera pose
    looking into the field, and spp drawn from `spp_choices`. Returns (scenes, views) with views as
    dicts of scene index, width, height, spp, frame and camera."""
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(num_scenes):
        tris = math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))
        n = max(1, int(round(tris / 1280)))
        box = 100.0 * (n / 1024.0) ** (1.0 / 3.0)
        scenes.append(sphere_field(seed * 1000 + k, num_spheres=n, num_emissive=min(4096, n * 1280 // 4), box=box,
                                   aspect=1.0))
    views = []
    for v in range(num_views):
        k = int(rng.integers(num_scenes))
        d = scenes[k]
        lo, hi = d.positions.min(0), d.positions.max(0)
        c = 0.5 * (lo + hi)
        ext = float((hi - lo).max())
        theta = rng.uniform(0, 2 * math.pi)
        eye = c + np.array([math.cos(theta), rng.uniform(-0.3, 0.3), math.sin(theta)]) * 0.9 * ext
        target = c + rng.normal(0, 0.05 * ext, 3)
        cam = Camera(tuple(eye), tuple(target), (0, 1, 0), math.radians(60.0), 1.0)
        views.append(dict(scene=k, width=res, height=res, spp=int(rng.choice(spp_choices)), frame=v, camera=cam))
    return scenes, views


def textured_spheres(seed: int = 3, num_objects: int = 64, tex_res: int = 512, subdivisions: int = 3,
                     target: bool = False, init_albedo: float = 0.5, init_roughness: float = 0.5) -> SceneData:
    """Differentiable-materials scene: an 8x8 grid of icospheres, each with its own RGB fp32 albedo
    texture and roughness, lit by an emissive quad. `target=True` gives the seeded ground truth
    (smooth random textures, roughness U[0.05, 1]); otherwise albedo `init_albedo` and roughness
    `init_roughness` everywhere."""
    rng = np.random.default_rng(seed)
    ico = TriangleMesh.create_icosphere(1.0, subdivisions)
    sb = SceneBuilder()
    g = int(math.ceil(math.sqrt(num_objects)))
    for i in range(num_objects):
        if target:
            tex = _smooth_texture(rng, tex_res)
            rough = float(rng.uniform(0.05, 1.0))
        else:
            tex = np.full((tex_res, tex_res, 3), init_albedo, np.float32)
            rough = init_roughness
        ti = sb.add_texture(tex)
        m = sb.add_material(standard_material(base=(1, 1, 1), roughness=rough, metallic=0.0, texture=ti))
        gx, gy = i % g, i // g
        t = Transform((1.0, 1.0, 1.0), (2.5 * (gx - (g - 1) / 2), 2.5 * (gy - (g - 1) / 2), 0.0),
                      rotation_matrix=random_rotation(np.random.default_rng(1000 + i)))
        sb.add_mesh_instance(ico, m, t, f"object{i}")
    light = sb.add_material(standard_material(base=(0, 0, 0), roughness=1.0, emissive=(4, 4, 4)))
    sb.add_mesh_instance(TriangleMesh.create_quad(), light, Transform.deg((40, 1, 40), (0, 0, 30), (-90, 0, 0)), "Light")
    sb.camera = Camera((0, 0, 24), (0, 0, 0), (0, 1, 0), math.radians(45.0), 1.0)
    return sb.build()


def _smooth_texture(rng: np.random.Generator, res: int, octaves: int = 4) -> np.ndarray:
    """Seeded smooth-noise RGB texture with values in [0.1, 0.9]."""
    out = np.zeros((res, res, 3), np.float64)
    amp = 1.0
    for o in range(octaves):
        n = 4 * 2 ** o
        coarse = rng.random((n, n, 3))
        idx = np.arange(res) * n / res
        i0 = np.floor(idx).astype(int) % n
        i1 = (i0 + 1) % n
        f = (idx - np.floor(idx))[:, None]
        rows = coarse[i0] * (1 - f[:, :, None]) + coarse[i1] * f[:, :, None]
        cols = rows[:, i0] * (1 - f[None, :, :]) + rows[:, i1] * f[None, :, :]
        out += amp * cols
        amp *= 0.5
    out = (out - out.min()) / max(out.max() - out.min(), 1e-9)
    return (0.1 + 0.8 * out).astype(np.float32)
