"""Test infrastructure: a plain-Python restatement of the scene ingestion
(OBJ / MTL triangle and material tables, .pbrdf parameters), written
independently of the product's C++ reader and used only by tests/ to check
it.

Follows add_objfile_to_flat_tri_list (src/hl-main.cpp:277-350),
convert_and_upload_materials (src/material-conversion.cpp:59-151) and
PrincipledBRDFParameters::readFrom (src/principledParameters.h:64-125), with
the reader conventions the product states for the external OBJ library
(fan triangulation, geometric normals when a face has none, relative
negative indices, undefined usemtl -> fallback).  Arithmetic is FP64 and
rounded to FP32 once per stored value, as the product does.
"""
from __future__ import annotations

import math
import os

import numpy as np

FALLBACK = "gi/fallback"
PBRDF_KEYS = {"color": ("color", 3), "metallic": ("metallic", 1), "opacity": ("opacity", 1),
              "subsurface": ("subsurface", 1), "specular": ("specular", 1), "roughness": ("roughness", 1),
              "tintSpecular": ("specular_tint", 1), "anisotropic": ("anisotropic", 1), "sheen": ("sheen", 1),
              "tintSheen": ("sheen_tint", 1), "clearcoat": ("clearcoat", 1),
              "glossClearcoat": ("clearcoat_gloss", 1), "transmission": ("transmission", 1)}


def read_mtl(path):
    out, cur = {}, None
    for line in open(path):
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        if tok[0] == "newmtl":
            cur = line.strip()[len("newmtl"):].strip()
            out[cur] = {"kd": [0.0] * 3, "ks": [0.0] * 3, "alpha": 1.0, "tex_d": False}
        elif cur is None:
            continue
        elif tok[0] == "Kd":
            out[cur]["kd"] = [float(x) for x in tok[1:4]]
        elif tok[0] == "Ks":
            out[cur]["ks"] = [float(x) for x in tok[1:4]]
        elif tok[0] == "d":
            out[cur]["alpha"] = float(tok[1])
        elif tok[0] == "Tr":
            out[cur]["alpha"] = 1.0 - float(tok[1])
        elif tok[0] == "map_Kd":
            out[cur]["tex_d"] = True
            if len(tok) > 1:  # the file is the last token, relative to the MTL's directory
                f = tok[-1]
                out[cur]["tex_path"] = f if f.startswith("/") else (os.path.dirname(path) or ".") + "/" + f
    return out


class ObjTable:
    def __init__(self):
        self.tris = []       # dicts: a,b,c,na,nb,nc,ta,tb,tc,mat
        self.mats = []       # dicts: kd, ks, alpha, principled params
        self.names = []
        self.tex = []
        self.textures = []   # distinct map_Kd files in order of first use

    def material(self, name, entry):
        if name in self.names:
            return self.names.index(name)
        self.names.append(name)
        if entry is None:
            self.mats.append({"kd": [1.0, 0.0, 0.0], "ks": [1.0, 0.0, 0.0], "alpha": 1.0, "texture": -1})
            self.tex.append(False)
        else:
            t = -1
            if entry["tex_d"] and entry.get("tex_path"):
                if entry["tex_path"] not in self.textures:
                    self.textures.append(entry["tex_path"])
                t = self.textures.index(entry["tex_path"])
            self.mats.append({"kd": list(entry["kd"]), "ks": list(entry["ks"]), "alpha": entry["alpha"],
                              "texture": t})
            self.tex.append(entry["tex_d"])
        return len(self.names) - 1

    def append(self, path, trafo=None):
        if not self.names:
            self.material(FALLBACK, None)
        M = None if trafo is None else np.asarray(trafo, np.float64).reshape(4, 4)
        NM = None if M is None else np.linalg.inv(M[:3, :3]).T
        d = os.path.dirname(path) or "."
        mtl, v, vn, vt = {}, [], [], []
        cur, cur_tex = self.names.index(FALLBACK), False
        for line in open(path):
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            k = tok[0]
            if k == "v":
                p = [float(x) for x in tok[1:4]]
                if M is not None:
                    q = M @ np.array(p + [1.0])
                    p = list(q[:3] / q[3])
                v.append(p)
            elif k == "vn":
                n = [float(x) for x in tok[1:4]]
                if NM is not None:
                    q = NM @ np.array(n)
                    ln = math.sqrt(float(q @ q))
                    n = list(q / ln) if ln > 0 else list(q)
                vn.append(n)
            elif k == "vt":
                vt.append([float(tok[1]), float(tok[2])])
            elif k == "mtllib":
                for lib in tok[1:]:
                    mtl.update(read_mtl(os.path.join(d, lib)))
            elif k == "usemtl":
                name = line.strip()[len("usemtl"):].strip()
                if name in mtl:
                    cur = self.material(name, mtl[name])
                    cur_tex = self.tex[cur]
                else:
                    cur, cur_tex = self.names.index(FALLBACK), False
            elif k == "f":
                corners = []
                for t in tok[1:]:
                    parts = t.split("/")
                    idx = []
                    for j, lst in enumerate((v, vt, vn)):
                        if j < len(parts) and parts[j] != "":
                            x = int(parts[j])
                            idx.append(x - 1 if x > 0 else len(lst) + x)
                        else:
                            idx.append(None)
                    corners.append(idx)
                for i in range(1, len(corners) - 1):
                    cs = (corners[0], corners[i], corners[i + 1])
                    P = [np.float32(v[c[0]]) for c in cs]
                    if all(c[2] is not None for c in cs):
                        N = [np.float32(vn[c[2]]) for c in cs]
                    else:
                        a, b, c = (np.asarray(p, np.float64) for p in P)
                        g = np.cross(b - a, c - a)
                        ln = math.sqrt(float(g @ g))
                        if ln > 0:
                            g = g / ln
                        N = [np.float32(g)] * 3
                    T = [np.float32(vt[c[1]]) if (cur_tex and c[1] is not None) else np.zeros(2, np.float32)
                         for c in cs]
                    self.tris.append({"a": P[0], "b": P[1], "c": P[2], "na": N[0], "nb": N[1], "nc": N[2],
                                      "ta": T[0], "tb": T[1], "tc": T[2], "mat": cur})

    def apply_pbrdf(self, directory):
        for i, name in enumerate(self.names):
            p = os.path.join(directory, name + ".pbrdf")
            if os.path.exists(p):
                self.mats[i]["pbrdf"] = read_pbrdf(p, self.mats[i]["kd"])


def read_pbrdf(path, kd=(1.0, 1.0, 1.0)):
    """Parameters of one .pbrdf file: defaults of the default constructor
    (opacity 1, transmission 1, others 0), color defaulting to kd."""
    r = {"metallic": 0.0, "subsurface": 0.0, "specular": 0.0, "roughness": 0.0, "specular_tint": 0.0,
         "anisotropic": 0.0, "sheen": 0.0, "sheen_tint": 0.0, "clearcoat": 0.0, "clearcoat_gloss": 0.0,
         "opacity": 1.0, "transmission": 1.0, "color": list(kd)}
    toks = open(path).read().split()
    i = 0
    while i < len(toks):
        key = toks[i]
        i += 1
        if key in PBRDF_KEYS:
            field, n = PBRDF_KEYS[key]
            vals = [float(x) for x in toks[i:i + n]]
            i += n
            r[field] = vals if n == 3 else vals[0]
    return r
