"""Loader for Falcor `.pyscene` scene descriptions (the subset this renderer models).

A `.pyscene` file is a Python script run against Falcor's scene-builder bindings
(Source/Falcor/Scene/Importers/PythonImporter.cpp; examples Scene/Box/CornellBox.pyscene,
Scene/Sphere/Sphere.pyscene). This loader never executes it: the file is parsed with `ast` and a
small interpreter evaluates a whitelist of statements (assignments, attribute assignments,
expression calls, `for` over `range`/lists, `if`) and expressions (literals, names, attribute
access, calls of the objects below, keyword arguments, arithmetic, tuples and lists). Anything
else raises `PySceneError` with the line number.

Objects available to a scene script, after Falcor's bindings:
  float2/float3/float4; StandardMaterial(name) and PBRTDiffuseMaterial(name) with baseColor,
  roughness, metallic, emissiveColor, emissiveFactor, indexOfRefraction, doubleSided,
  specularTransmission, diffuseTransmission, transmissionColor, thinSurface and
  loadTexture(MaterialTextureSlot.BaseColor, path, useSrgb) (StandardMaterial.cpp:261-268,
  BasicMaterial.cpp:681-691, Material.cpp:391-405; defaults of BasicMaterialData.slang:50-57,
  base colour, specular and transmission parameters and the IoR stored in float16 as the
  reference stores them);
  TriangleMesh.createQuad/createCube/createSphere/createDisk/createFromFile (.obj)
  (TriangleMesh.cpp:325-363); Transform(translation, scaling, rotationEuler, rotationEulerDeg)
  (Transform.cpp:151-213); Camera with position, target, up, focalLength, frameHeight,
  focalDistance, nearPlane, farPlane, aspectRatio (Camera.cpp:448-458; fovY from the focal
  length and frame height); sceneBuilder.addNode/addTriangleMesh/addMeshInstance/addCamera.
EnvMap(path) with intensity, tint and rotation, assigned to sceneBuilder.envMap, becomes the
scene's environment map. PointLight (intensity, position, direction, openingAngle,
penumbraAngle) and DirectionalLight (intensity, direction) passed to sceneBuilder.addLight become
analytic lights; other light types are accepted and ignored with a warning.
"""
from __future__ import annotations

import ast
import math
import operator
from pathlib import Path

import numpy as np

from . import scene as S


class PySceneError(ValueError):
    pass


# --------------------------------------------------------------------------- value types
class Vec:
    """float2/3/4 value (element-wise arithmetic, .x .y .z .w)."""

    def __init__(self, *v):
        self.v = tuple(float(x) for x in v)

    @classmethod
    def make(cls, n):
        def ctor(*args):
            if len(args) == 1 and not isinstance(args[0], Vec):
                return cls(*([args[0]] * n))
            flat = []
            for a in args:
                flat.extend(a.v if isinstance(a, Vec) else [a])
            if len(flat) != n:
                raise PySceneError(f"float{n} needs {n} components")
            return cls(*flat)
        return ctor

    def _op(self, o, f):
        ov = o.v if isinstance(o, Vec) else (float(o),) * len(self.v)
        return Vec(*(f(a, b) for a, b in zip(self.v, ov)))

    __add__ = lambda s, o: s._op(o, operator.add)
    __radd__ = __add__
    __sub__ = lambda s, o: s._op(o, operator.sub)
    __rsub__ = lambda s, o: s._op(o, lambda a, b: b - a)
    __mul__ = lambda s, o: s._op(o, operator.mul)
    __rmul__ = __mul__
    __truediv__ = lambda s, o: s._op(o, operator.truediv)
    __neg__ = lambda s: Vec(*(-a for a in s.v))

    def __getattr__(self, k):
        i = "xyzw".find(k)
        if len(k) == 1 and 0 <= i < len(self.v):
            return self.v[i]
        raise AttributeError(k)

    def __len__(self):
        return len(self.v)

    def __getitem__(self, i):
        return self.v[i]


class MaterialTextureSlot:
    BaseColor, Specular, Emissive, Normal, Transmission, Displacement = range(6)


def _half(x):
    return float(np.float16(x))


class _Material:
    kind = S.MAT_STANDARD

    def __init__(self, name="", model=None):
        self.name = name
        self.baseColor = Vec(1, 1, 1, 1)
        self.roughness = 0.0
        self.metallic = 0.0
        self.emissiveColor = Vec(0, 0, 0)
        self.emissiveFactor = 1.0
        self.indexOfRefraction = 1.5
        self.doubleSided = False
        self.specularTransmission = 0.0
        self.diffuseTransmission = 0.0
        self.transmissionColor = Vec(1, 1, 1)
        self.thinSurface = False
        self.texture_path = None
        self.texture_srgb = True
        self.emissive_path = None
        self.emissive_srgb = True
        self._warnings = []

    def loadTexture(self, slot, path, useSrgb=True):
        if slot == MaterialTextureSlot.BaseColor:
            self.texture_path, self.texture_srgb = str(path), bool(useSrgb)
        elif slot == MaterialTextureSlot.Emissive:
            self.emissive_path, self.emissive_srgb = str(path), bool(useSrgb)
        else:
            self._warnings.append(f"material '{self.name}': texture slot {slot} ignored (base colour and emissive only)")
        return True

    load_texture = loadTexture

    def record(self, texture_index=-1, emissive_index=-1) -> np.ndarray:
        bc = self.baseColor.v
        if self.kind == S.MAT_PBRT_DIFFUSE:
            return S.pbrt_diffuse_material([_half(c) for c in bc[:3]], self.doubleSided, texture_index)
        e = self.emissiveColor.v if isinstance(self.emissiveColor, Vec) else (self.emissiveColor,) * 3
        t = self.transmissionColor.v if isinstance(self.transmissionColor, Vec) else (self.transmissionColor,) * 3
        return S.standard_material([_half(c) for c in bc[:3]], _half(self.roughness), _half(self.metallic),
                                   _half(self.indexOfRefraction), e[:3], self.emissiveFactor, self.doubleSided,
                                   texture_index, emissive_index, _half(self.specularTransmission),
                                   _half(self.diffuseTransmission), [_half(c) for c in t[:3]], bool(self.thinSurface))


class StandardMaterial(_Material):
    kind = S.MAT_STANDARD


class PBRTDiffuseMaterial(_Material):
    kind = S.MAT_PBRT_DIFFUSE


class _Transform:
    def __init__(self, **kw):
        self.translation = Vec(0, 0, 0)
        self.scaling = Vec(1, 1, 1)
        self.rotation = (0.0, 0.0, 0.0)
        for k, v in kw.items():
            if k == "translation":
                self.translation = v
            elif k == "scaling":
                self.scaling = v if isinstance(v, Vec) else Vec(v, v, v)
            elif k == "rotationEuler":
                self.rotation = tuple(v.v)
            elif k == "rotationEulerDeg":
                self.rotation = tuple(math.radians(a) for a in v.v)
            else:
                raise PySceneError(f"Transform argument '{k}' is not supported")

    def to_scene(self) -> S.Transform:
        return S.Transform(tuple(self.scaling.v), tuple(self.translation.v), self.rotation)


class _Camera:
    def __init__(self, name=""):
        self.name = name
        self.position = Vec(0, 0, 0)
        self.target = Vec(0, 0, -1)
        self.up = Vec(0, 1, 0)
        self.focalLength = 21.0      # CameraData.slang defaults
        self.frameHeight = 24.0
        self.focalDistance = 10000.0
        self.nearPlane = 0.1
        self.farPlane = 1000.0
        self.aspectRatio = None
        self.apertureRadius = 0.0

    def to_scene(self, aspect) -> S.Camera:
        fov = 2.0 * math.atan(0.5 * self.frameHeight / self.focalLength)
        return S.Camera(tuple(self.position.v), tuple(self.target.v), tuple(self.up.v), fov,
                        self.aspectRatio if self.aspectRatio else aspect, self.focalDistance, self.nearPlane, self.farPlane)


class _Ignored:
    """Analytic lights, environment maps and other objects this renderer does not model."""

    def __init__(self, kind, *a, **kw):
        self.kind = kind

    def __setattr__(self, k, v):
        object.__setattr__(self, k, v)


class _EnvMap:
    """EnvMap(path) with intensity, tint and rotation (degrees XYZ) (EnvMap.cpp:140-160)."""

    def __init__(self, path):
        self.path = str(path)
        self.intensity = 1.0
        self.tint = Vec(1, 1, 1)
        self.rotation = Vec(0, 0, 0)


class _TriangleMeshAPI:
    def __init__(self, base: Path, warnings: list):
        self.base = base
        self.warnings = warnings

    @staticmethod
    def createQuad(size=None):
        s = size.v if isinstance(size, Vec) else (1.0, 1.0)
        return S.TriangleMesh.create_quad(s)

    @staticmethod
    def createCube(size=None):
        s = size.v if isinstance(size, Vec) else (1.0, 1.0, 1.0)
        return S.TriangleMesh.create_cube(s)

    @staticmethod
    def createSphere(radius=1.0, segmentsU=32, segmentsV=32):
        return S.TriangleMesh.create_sphere(radius, int(segmentsU), int(segmentsV))

    @staticmethod
    def createDisk(radius=1.0, segments=32):
        # Triangle fan in the XZ plane facing +Y (TriangleMesh.cpp:80-100).
        n = int(segments)
        ang = np.arange(n) * (2 * np.pi / n)
        p = np.concatenate([[[0, 0, 0]], np.stack([np.cos(ang), np.zeros(n), np.sin(ang)], 1) * radius]).astype(np.float32)
        nrm = np.tile(np.array([[0, 1, 0]], np.float32), (n + 1, 1))
        uv = np.concatenate([[[0.5, 0.5]], np.stack([np.cos(ang), np.sin(ang)], 1) * 0.5 + 0.5]).astype(np.float32)
        idx = np.array([[0, 1 + (i + 1) % n, 1 + i] for i in range(n)], np.uint32)
        return S.TriangleMesh(p, nrm, uv, idx)

    def createFromFile(self, path, smoothNormals=False, importFlags=None):
        return load_obj(self.base / str(path), smooth_normals=bool(smoothNormals))


class _AnalyticLight:
    """PointLight / DirectionalLight with the reference's properties (Light.cpp:408-428: intensity,
    position, direction, openingAngle, penumbraAngle; LightData.slang defaults)."""

    def __init__(self, name=""):
        self.name = str(name)
        self.intensity = Vec(1, 1, 1)
        self.direction = Vec(0, -1, 0)
        self.active = True

    def _v(self, x):
        return tuple(x.v) if isinstance(x, Vec) else (float(x),) * 3


class PointLight(_AnalyticLight):
    kind = "PointLight"

    def __init__(self, name=""):
        super().__init__(name)
        self.position = Vec(0, 0, 0)
        self.openingAngle = math.pi
        self.penumbraAngle = 0.0

    def record(self):
        return S.point_light(self._v(self.intensity), self._v(self.position), self._v(self.direction),
                             float(self.openingAngle), float(self.penumbraAngle))


class DirectionalLight(_AnalyticLight):
    kind = "DirectionalLight"

    def record(self):
        return S.directional_light(self._v(self.intensity), self._v(self.direction))


class _SceneBuilderAPI:
    def __init__(self, sb: S.SceneBuilder, base: Path, warnings: list):
        self.sb = sb
        self.base = base
        self.warnings = warnings
        self.nodes = []
        self.meshes = []
        self.cameras = []
        self.materials = {}
        self.textures = {}
        self.envMap = None

    def addNode(self, name, transform=None, parent=None):
        if parent is not None:
            raise PySceneError("node hierarchies (parent) are not supported")
        self.nodes.append((str(name), transform or _Transform()))
        return len(self.nodes) - 1

    def addTriangleMesh(self, mesh, material):
        self.meshes.append((mesh, material))
        return len(self.meshes) - 1

    def _material_id(self, m: _Material) -> int:
        key = id(m)
        if key not in self.materials:
            def texture(path, srgb):
                if not path:
                    return -1
                tp = self.base / path
                if (tp, srgb) not in self.textures:
                    self.textures[(tp, srgb)] = self.sb.add_texture(_load_texture(tp, srgb))
                return self.textures[(tp, srgb)]
            tex = texture(m.texture_path, m.texture_srgb)
            etex = texture(m.emissive_path, m.emissive_srgb) if m.kind == S.MAT_STANDARD else -1
            self.materials[key] = self.sb.add_material(m.record(tex, etex))
            self.warnings.extend(m._warnings)
        return self.materials[key]

    def addMeshInstance(self, node_id, mesh_id):
        name, t = self.nodes[node_id]
        mesh, mat = self.meshes[mesh_id]
        self.sb.add_mesh_instance(mesh, self._material_id(mat), t.to_scene(), name)

    def addCamera(self, camera):
        self.cameras.append(camera)
        return len(self.cameras) - 1

    def addLight(self, light):
        if isinstance(light, _AnalyticLight):
            return self.sb.add_light(light.record())
        self.warnings.append(f"{getattr(light, 'kind', 'light')} ignored (point and directional lights are supported)")
        return 0

    def __setattr__(self, k, v):
        if k == "envMap" and v is not None and not isinstance(v, _EnvMap):
            raise PySceneError("sceneBuilder.envMap must be an EnvMap")
        object.__setattr__(self, k, v)


def _srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def _load_texture(path: Path, srgb: bool) -> np.ndarray:
    from .image_io import read_rgba
    rgba = read_rgba(path)
    rgb = rgba[..., :3]
    if srgb and path.suffix.lower() not in (".exr", ".pfm", ".hdr"):
        rgb = _srgb_to_linear(rgb)
    return np.ascontiguousarray(rgb, np.float32)


def load_obj(path, smooth_normals=False) -> S.TriangleMesh:
    """Wavefront OBJ (v, vt, vn, f with polygons fanned into triangles; negative indices)."""
    pos, uvs, nrms = [], [], []
    verts = {}
    out_p, out_n, out_t, idx = [], [], [], []
    have_n = True
    for line in Path(path).read_text().splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "v":
            pos.append([float(x) for x in f[1:4]])
        elif f[0] == "vt":
            uvs.append([float(x) for x in f[1:3]])
        elif f[0] == "vn":
            nrms.append([float(x) for x in f[1:4]])
        elif f[0] == "f":
            corner = []
            for c in f[1:]:
                parts = (c.split("/") + ["", ""])[:3]
                ids = []
                for j, (s, n) in enumerate(zip(parts, (len(pos), len(uvs), len(nrms)))):
                    ids.append(None if s == "" else (int(s) - 1 if int(s) > 0 else n + int(s)))
                if ids[2] is None:
                    have_n = False
                key = tuple(ids)
                if key not in verts:
                    verts[key] = len(out_p)
                    out_p.append(pos[ids[0]])
                    out_t.append(uvs[ids[1]] if ids[1] is not None else [0.0, 0.0])
                    out_n.append(nrms[ids[2]] if ids[2] is not None else [0.0, 0.0, 0.0])
                corner.append(verts[key])
            for k in range(1, len(corner) - 1):
                idx.append([corner[0], corner[k], corner[k + 1]])
    p = np.asarray(out_p, np.float32)
    i = np.asarray(idx, np.uint32).reshape(-1, 3)
    n = S.vertex_normals(p, i) if (smooth_normals or not have_n) else np.asarray(out_n, np.float32)
    return S.TriangleMesh(p, np.asarray(n, np.float32), np.asarray(out_t, np.float32), i)


# --------------------------------------------------------------------------- interpreter
_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.FloorDiv: operator.floordiv, ast.Mod: operator.mod, ast.Pow: operator.pow}
_CMPOPS = {ast.Eq: operator.eq, ast.NotEq: operator.ne, ast.Lt: operator.lt, ast.LtE: operator.le, ast.Gt: operator.gt,
           ast.GtE: operator.ge}
_MAX_LOOP = 1_000_000


class _Interp:
    def __init__(self, env: dict):
        self.env = env
        self.steps = 0

    def fail(self, node, msg):
        raise PySceneError(f"line {getattr(node, 'lineno', '?')}: {msg}")

    def expr(self, n):
        if isinstance(n, ast.Constant):
            if isinstance(n.value, (int, float, str, bool)) or n.value is None:
                return n.value
            self.fail(n, "unsupported literal")
        if isinstance(n, ast.Name):
            if n.id not in self.env:
                self.fail(n, f"unknown name '{n.id}'")
            return self.env[n.id]
        if isinstance(n, ast.Attribute):
            if n.attr.startswith("_"):
                self.fail(n, "private attribute")
            obj = self.expr(n.value)
            if isinstance(obj, (int, float, str, bool, list, tuple)):
                self.fail(n, f"attribute '{n.attr}' of a {type(obj).__name__}")
            try:
                return getattr(obj, n.attr)
            except AttributeError:
                self.fail(n, f"unknown attribute '{n.attr}'")
        if isinstance(n, ast.Call):
            fn = self.expr(n.func)
            if not callable(fn) or not getattr(fn, "_pyscene_ok", True) or fn in (getattr, setattr, eval, exec):
                self.fail(n, "call of a non-callable")
            args = [self.expr(a) for a in n.args]
            kw = {k.arg: self.expr(k.value) for k in n.keywords if k.arg is not None}
            if len(kw) != len(n.keywords):
                self.fail(n, "** arguments are not supported")
            try:
                return fn(*args, **kw)
            except PySceneError as e:
                if str(e).startswith("line "):
                    raise
                self.fail(n, str(e))
            except (TypeError, ValueError, KeyError, IndexError, OSError) as e:
                self.fail(n, str(e))
        if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
            return _BINOPS[type(n.op)](self.expr(n.left), self.expr(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd, ast.Not)):
            v = self.expr(n.operand)
            return -v if isinstance(n.op, ast.USub) else (v if isinstance(n.op, ast.UAdd) else not v)
        if isinstance(n, (ast.Tuple, ast.List)):
            vals = [self.expr(e) for e in n.elts]
            return tuple(vals) if isinstance(n, ast.Tuple) else vals
        if isinstance(n, ast.Compare) and len(n.ops) == 1 and type(n.ops[0]) in _CMPOPS:
            return _CMPOPS[type(n.ops[0])](self.expr(n.left), self.expr(n.comparators[0]))
        if isinstance(n, ast.IfExp):
            return self.expr(n.body) if self.expr(n.test) else self.expr(n.orelse)
        if isinstance(n, ast.Subscript):
            return self.expr(n.value)[self.expr(n.slice)]
        self.fail(n, f"unsupported expression {type(n).__name__}")

    def assign(self, target, value, node):
        if isinstance(target, ast.Name):
            self.env[target.id] = value
        elif isinstance(target, ast.Attribute):
            if target.attr.startswith("_"):
                self.fail(node, "private attribute")
            obj = self.expr(target.value)
            if isinstance(obj, _SceneBuilderAPI):
                ok = target.attr == "envMap"
            elif isinstance(obj, (_Material, _Camera, _Transform, _EnvMap, _AnalyticLight)):
                # Only the existing data attributes of these bindings.
                ok = hasattr(obj, target.attr) and not callable(getattr(obj, target.attr))
            else:
                ok = isinstance(obj, _Ignored)
            if not ok:
                self.fail(node, f"cannot set '{target.attr}' of {type(obj).__name__.strip('_')}")
            setattr(obj, target.attr, value)
        elif isinstance(target, (ast.Tuple, ast.List)):
            vals = list(value)
            if len(vals) != len(target.elts):
                self.fail(node, "unpacking mismatch")
            for t, v in zip(target.elts, vals):
                self.assign(t, v, node)
        else:
            self.fail(node, "unsupported assignment target")

    def run(self, body):
        for st in body:
            self.steps += 1
            if self.steps > _MAX_LOOP:
                self.fail(st, "scene script too long")
            if isinstance(st, ast.Assign):
                v = self.expr(st.value)
                for t in st.targets:
                    self.assign(t, v, st)
            elif isinstance(st, ast.AugAssign) and type(st.op) in _BINOPS:
                cur = self.expr(ast.Name(id=st.target.id, ctx=ast.Load(), lineno=st.lineno)) if isinstance(
                    st.target, ast.Name) else self.expr(st.target)
                self.assign(st.target, _BINOPS[type(st.op)](cur, self.expr(st.value)), st)
            elif isinstance(st, ast.Expr):
                if isinstance(st.value, ast.Constant):
                    continue  # docstrings
                self.expr(st.value)
            elif isinstance(st, ast.For):
                it = self.expr(st.iter)
                if not isinstance(it, (range, list, tuple)):
                    self.fail(st, "for loops run over range(), lists or tuples")
                if len(it) > _MAX_LOOP:
                    self.fail(st, "loop too long")
                for v in it:
                    self.assign(st.target, v, st)
                    self.run(st.body)
            elif isinstance(st, ast.If):
                self.run(st.body if self.expr(st.test) else st.orelse)
            elif isinstance(st, ast.Pass):
                pass
            elif isinstance(st, (ast.Import, ast.ImportFrom)):
                names = [a.name for a in st.names]
                if all(n in ("falcor", "*", "math") for n in names) or getattr(st, "module", None) == "falcor":
                    continue  # the bindings are already in scope
                self.fail(st, f"import of {names} is not supported")
            else:
                self.fail(st, f"unsupported statement {type(st).__name__}")


def load_pyscene(path, aspect: float = 1.0, warnings: list | None = None) -> S.SceneData:
    """Builds SceneData from a .pyscene file (see the module docstring for the supported subset).
    aspect: the camera's aspect ratio unless the script sets camera.aspectRatio. Ignored calls
    and unsupported material features are appended to `warnings`."""
    path = Path(path)
    base = path.parent
    warns = warnings if warnings is not None else []
    tree = ast.parse(path.read_text(), filename=str(path))
    sb = S.SceneBuilder()
    api = _SceneBuilderAPI(sb, base, warns)
    env = {
        "float2": Vec.make(2), "float3": Vec.make(3), "float4": Vec.make(4),
        "StandardMaterial": StandardMaterial, "PBRTDiffuseMaterial": PBRTDiffuseMaterial,
        "MaterialTextureSlot": MaterialTextureSlot, "TriangleMesh": _TriangleMeshAPI(base, warns),
        "Transform": _Transform, "Camera": _Camera, "sceneBuilder": api,
        "range": range, "len": len, "abs": abs, "min": min, "max": max, "float": float, "int": int, "str": str,
        "math": type("math", (), {"pi": math.pi, "radians": staticmethod(math.radians), "sin": staticmethod(math.sin),
                                  "cos": staticmethod(math.cos), "sqrt": staticmethod(math.sqrt)}),
        "True": True, "False": False, "None": None,
    }
    env["EnvMap"] = _EnvMap
    env["PointLight"] = PointLight
    env["DirectionalLight"] = DirectionalLight
    for kind in ("DistantLight", "RectLight", "DiscLight", "SphereLight",
                 "GridVolume", "Grid", "SDFGrid", "SDFSVS", "SDFSBS", "SDFSVO", "NDSDFGrid"):
        env[kind] = (lambda k: (lambda *a, **kw: _Ignored(k, *a, **kw)))(kind)
    _Interp(env).run(tree.body)
    if not sb.materials:
        raise PySceneError(f"{path}: the scene has no mesh instances")
    cam = api.cameras[0] if api.cameras else _Camera()
    if not api.cameras:
        warns.append("no camera in the scene: default camera")
    sb.camera = cam.to_scene(aspect)
    data = sb.build()
    if api.envMap is not None:
        from .image_io import read_rgba
        e = api.envMap
        tint = e.tint.v if isinstance(e.tint, Vec) else (float(e.tint),) * 3
        rot = e.rotation.v if isinstance(e.rotation, Vec) else (0.0, 0.0, 0.0)
        data.env_map = {"texels": read_rgba(base / e.path)[..., :3], "intensity": float(e.intensity),
                        "tint": tuple(tint[:3]), "rotation": tuple(rot[:3])}
    return data
