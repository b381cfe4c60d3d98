"""Scene ingestion: the `caustic-scene` v1 JSON format of serialize.ts
(serialize.ts:8-179) with the same validation rules and error messages.

The reference's plain objects ({x, y, z}, {r, g, b}) map to this package's
scene dicts (tuples; mesh vertices as an (n, 3) array), which packScene takes.
"""
import json
import math

import numpy as np

SCENE_FORMAT = "caustic-scene"
SCENE_FORMAT_VERSION = 1
MAX_OBJECTS = 20000


class SceneParseError(ValueError):
    """Raised for any invalid scene document (serialize.ts:11-16)."""


def _fail(msg):
    raise SceneParseError(msg)


def _is_obj(v):
    return isinstance(v, dict)


def _fmt(x):
    """A bound as JS prints it inside a template string."""
    if x == math.inf:
        return "Infinity"
    if x == -math.inf:
        return "-Infinity"
    if float(x).is_integer():
        return str(int(x))
    r = repr(float(x))
    return np.format_float_positional(float(x), trim="-") if "e" in r else r


def _js_string(v):
    """String(v) as JS prints a JSON value (null, true, numbers, arrays, objects)."""
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, float)):
        try:
            return _fmt(float(v))
        except OverflowError:
            return "Infinity" if v > 0 else "-Infinity"
    if isinstance(v, list):
        return ",".join("" if x is None else _js_string(x) for x in v)
    if isinstance(v, dict):
        return "[object Object]"
    return str(v)


def _num(v, path, lo=-math.inf, hi=math.inf):
    # JSON numbers are binary64 in JS: an integer literal beyond the double
    # range is Infinity there, so it fails as non-finite here too
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        _fail(f"{path} must be a finite number")
    try:
        v = float(v)
    except OverflowError:
        _fail(f"{path} must be a finite number")
    if not math.isfinite(v):
        _fail(f"{path} must be a finite number")
    if v < lo or v > hi:
        _fail(f"{path} must be within [{_fmt(lo)}, {_fmt(hi)}]")
    return v


def _str(v, path):
    if not isinstance(v, str):
        _fail(f"{path} must be a string")
    return v


def _vec(v, path):
    if not _is_obj(v):
        _fail(f"{path} must be an object {{x, y, z}}")
    return (_num(v.get("x"), f"{path}.x"), _num(v.get("y"), f"{path}.y"),
            _num(v.get("z"), f"{path}.z"))


def _color(v, path, hi=1e4):
    if not _is_obj(v):
        _fail(f"{path} must be an object {{r, g, b}}")
    return (_num(v.get("r"), f"{path}.r", 0, hi), _num(v.get("g"), f"{path}.g", 0, hi),
            _num(v.get("b"), f"{path}.b", 0, hi))


def _material(v, path):
    if not _is_obj(v):
        _fail(f"{path} must be an object")
    t = v.get("type")
    if t == "lambertian":
        m = {"type": "lambertian", "albedo": _color(v.get("albedo"), f"{path}.albedo", 1)}
        if "checker" in v:
            c = v["checker"]
            if not _is_obj(c):
                _fail(f"{path}.checker must be an object")
            m["checker"] = {"color2": _color(c.get("color2"), f"{path}.checker.color2", 1),
                            "scale": _num(c.get("scale"), f"{path}.checker.scale", 1e-3, 1e4)}
        if "emission" in v:
            m["emission"] = _color(v["emission"], f"{path}.emission")
        return m
    if t == "metal":
        return {"type": "metal", "albedo": _color(v.get("albedo"), f"{path}.albedo", 1),
                "roughness": _num(v.get("roughness"), f"{path}.roughness", 0, 1)}
    if t == "dielectric":
        ior = _num(v.get("ior"), f"{path}.ior", 1, 4)
        if "tint" in v:
            return {"type": "dielectric", "ior": ior, "tint": _color(v["tint"], f"{path}.tint", 1)}
        return {"type": "dielectric", "ior": ior}
    if t == "emissive":
        return {"type": "emissive", "color": _color(v.get("color"), f"{path}.color", 1),
                "strength": _num(v.get("strength"), f"{path}.strength", 0, 1e5)}
    return _fail(f"{path}.type must be lambertian | metal | dielectric | emissive")


def _shape(v, path):
    if not _is_obj(v):
        _fail(f"{path} must be an object")
    k = v.get("kind")
    if k == "sphere":
        return {"kind": "sphere", "center": _vec(v.get("center"), f"{path}.center"),
                "radius": _num(v.get("radius"), f"{path}.radius", 1e-6, 1e6)}
    if k == "triangle":
        return {"kind": "triangle", "a": _vec(v.get("a"), f"{path}.a"),
                "b": _vec(v.get("b"), f"{path}.b"), "c": _vec(v.get("c"), f"{path}.c")}
    if k == "quad":
        return {"kind": "quad", "origin": _vec(v.get("origin"), f"{path}.origin"),
                "u": _vec(v.get("u"), f"{path}.u"), "v": _vec(v.get("v"), f"{path}.v")}
    if k == "mesh":
        if not isinstance(v.get("vertices"), list) or not isinstance(v.get("indices"), list):
            _fail(f"{path} mesh needs vertices[] and indices[]")
        verts = [_vec(p, f"{path}.vertices[{i}]") for i, p in enumerate(v["vertices"])]
        idx = [_num(n, f"{path}.indices[{i}]", 0, len(verts) - 1)
               for i, n in enumerate(v["indices"])]
        if len(idx) % 3 != 0:
            _fail(f"{path}.indices length must be a multiple of 3")
        if any(not float(n).is_integer() for n in idx):
            _fail(f"{path}.indices must be integers")
        return {"kind": "mesh", "vertices": np.array(verts, np.float64).reshape(-1, 3),
                "indices": np.array(idx, np.int64)}
    return _fail(f"{path}.kind must be sphere | triangle | quad | mesh")


def _camera(v, path):
    if not _is_obj(v):
        _fail(f"{path} must be an object")
    return {"target": _vec(v.get("target"), f"{path}.target"),
            "yaw": _num(v.get("yaw"), f"{path}.yaw"),
            "pitch": _num(v.get("pitch"), f"{path}.pitch", -math.pi / 2, math.pi / 2),
            "distance": _num(v.get("distance"), f"{path}.distance", 1e-3, 1e6),
            "fovDeg": _num(v.get("fovDeg"), f"{path}.fovDeg", 1, 170),
            "lensRadius": _num(v.get("lensRadius"), f"{path}.lensRadius", 0, 1e3),
            "focusDistance": _num(v.get("focusDistance"), f"{path}.focusDistance", 1e-3, 1e6)}


_MISSING = object()  # a key that is absent, as opposed to present with null


def _sky(v, path):
    if v is _MISSING:
        return {"kind": "none"}
    if not _is_obj(v):
        _fail(f"{path} must be an object")
    if v.get("kind") == "none":
        return {"kind": "none"}
    if v.get("kind") == "gradient":
        return {"kind": "gradient", "zenith": _color(v.get("zenith"), f"{path}.zenith"),
                "horizon": _color(v.get("horizon"), f"{path}.horizon"),
                "strength": _num(v.get("strength"), f"{path}.strength", 0, 1e3)}
    return _fail(f"{path}.kind must be none | gradient")


def parseScene(doc):
    """parseScene (serialize.ts:143-170): a validated scene dict from a parsed document."""
    if not _is_obj(doc):
        _fail("scene must be a JSON object")
    if "format" in doc and doc["format"] != SCENE_FORMAT:
        _fail(f'unknown format "{_js_string(doc["format"])}"')
    name = "Imported scene" if "name" not in doc else _str(doc["name"], "name")
    raw = doc.get("objects")
    if not isinstance(raw, list):
        _fail("objects must be an array")
    if len(raw) > MAX_OBJECTS:
        _fail(f"too many objects (max {MAX_OBJECTS})")
    seen, objects = set(), []
    for i, r in enumerate(raw):
        if not _is_obj(r):
            _fail(f"objects[{i}] must be an object")
        oid = f"obj-{i}" if "id" not in r else _str(r["id"], f"objects[{i}].id")
        while oid in seen:
            oid = f"{oid}-dup"
        seen.add(oid)
        objects.append({"id": oid,
                        "name": f"Object {i + 1}" if "name" not in r else
                        _str(r["name"], f"objects[{i}].name"),
                        "shape": _shape(r.get("shape"), f"objects[{i}].shape"),
                        "material": _material(r.get("material"), f"objects[{i}].material")})
    return {"name": name, "objects": objects, "camera": _camera(doc.get("camera"), "camera"),
            "sky": _sky(doc.get("sky", _MISSING), "sky")}


def parseSceneJSON(text):
    """parseSceneJSON (serialize.ts:172-179)."""
    try:
        doc = json.loads(text)
    except ValueError as e:
        raise SceneParseError(f"invalid JSON: {e}") from None
    return parseScene(doc)


def _jvec(v):
    return {"x": float(v[0]), "y": float(v[1]), "z": float(v[2])}


def _jcolor(c):
    return {"r": float(c[0]), "g": float(c[1]), "b": float(c[2])}


def _jmaterial(m):
    out = {"type": m["type"]}
    for k, v in m.items():
        if k == "type":
            continue
        if k in ("albedo", "emission", "tint", "color"):
            out[k] = _jcolor(v)
        elif k == "checker":
            out[k] = {"color2": _jcolor(v["color2"]), "scale": float(v["scale"])}
        else:
            out[k] = float(v)
    return out


def _jshape(s):
    k = s["kind"]
    if k == "sphere":
        return {"kind": k, "center": _jvec(s["center"]), "radius": float(s["radius"])}
    if k == "triangle":
        return {"kind": k, "a": _jvec(s["a"]), "b": _jvec(s["b"]), "c": _jvec(s["c"])}
    if k == "quad":
        return {"kind": k, "origin": _jvec(s["origin"]), "u": _jvec(s["u"]), "v": _jvec(s["v"])}
    verts = np.asarray(s["vertices"], np.float64).reshape(-1, 3)
    return {"kind": "mesh", "vertices": [_jvec(p) for p in verts],
            "indices": [int(i) for i in np.asarray(s["indices"]).reshape(-1)]}


def sceneToJSON(scene, pretty=True):
    """sceneToJSON (serialize.ts:128-139)."""
    cam = scene["camera"]
    doc = {"format": SCENE_FORMAT, "version": SCENE_FORMAT_VERSION, "name": scene.get("name", ""),
           "camera": {"target": _jvec(cam["target"]), **{k: float(cam[k]) for k in
                      ("yaw", "pitch", "distance", "fovDeg", "lensRadius", "focusDistance")}},
           "sky": ({"kind": "none"} if scene.get("sky", {"kind": "none"})["kind"] == "none" else
                   {"kind": "gradient", "zenith": _jcolor(scene["sky"]["zenith"]),
                    "horizon": _jcolor(scene["sky"]["horizon"]),
                    "strength": float(scene["sky"]["strength"])}),
           "objects": []}
    for i, o in enumerate(scene["objects"]):
        j = {}
        if "id" in o:
            j["id"] = o["id"]
        j["name"] = o.get("name", f"Object {i + 1}")
        j["shape"] = _jshape(o["shape"])
        j["material"] = _jmaterial(o["material"])
        doc["objects"].append(j)
    return json.dumps(doc, indent=2 if pretty else None)
