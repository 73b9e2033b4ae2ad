"""Shared test inputs: seeded glaciers and smooth velocity fields."""
import numpy as np
import torch

from igm_amd import synthetic as syn
from oracle import vert_disc as VD


def dzeta_for(nz, vs=4.0):
    return VD.compute_zeta_dzeta(VD.compute_levels(nz, vs))[1]


def small_glacier(ny, nx, dx=100.0, seed=0, A=None, C=None):
    """A dome on a rough inclined bed, sized to the grid (ny, nx)."""
    rng = np.random.default_rng(seed)
    bed = 2500.0 - 0.1 * np.arange(nx)[None, :] * dx + syn.grf((ny, nx), 20.0, 3 * dx, dx, seed)
    R = 0.45 * min(ny, nx) * dx
    H = syn.domes((ny, nx), dx, [((nx - 1) * dx / 2, (ny - 1) * dx / 2)], [R], [300.0])
    if A is None:
        A = rng.uniform(50, 100, (ny, nx))
    if C is None:
        C = 0.0464
    return syn._pack(bed, H, A, C, dx)


def smooth_velocity(nz, ny, nx, seed=0, scale=100.0):
    """A smooth, level-dependent (U, V) pair (float64 numpy), non-trivial shear."""
    rng = np.random.default_rng(seed)
    base_u = syn.grf((ny, nx), 1.0, 4.0, 1.0, seed + 1) if min(ny, nx) > 2 else rng.standard_normal((ny, nx))
    base_v = syn.grf((ny, nx), 1.0, 4.0, 1.0, seed + 2) if min(ny, nx) > 2 else rng.standard_normal((ny, nx))
    z = np.linspace(0, 1, nz)[:, None, None]
    U = scale * (0.5 + base_u[None]) * (0.3 + z ** 2) + 5 * rng.standard_normal((nz, ny, nx))
    V = scale * (0.2 + base_v[None]) * (0.3 + z) + 5 * rng.standard_normal((nz, ny, nx))
    return U, V


def to_dev(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def h5_structure(path):
    """Groups, attributes (type, shape, value) and datasets (type, shape) of an
    HDF5 file, from h5dump's XML dump of its header and attributes (a data
    dump: nothing in the file is executed).  String values are unescaped; a
    ``model_config`` attribute is parsed as JSON.  Dataset values are not read."""
    import json
    import re
    import subprocess
    import xml.etree.ElementTree as ET
    from igm_amd.keras_h5 import _h5dump
    xml = subprocess.run([_h5dump(), "-x", "-H", "-A", path], capture_output=True, text=True,
                         check=True).stdout
    H = "{http://hdfgroup.org/HDF5/XML/schema/HDF5-File.xsd}"

    def dtype(node):
        st, fl = node.find(f".//{H}StringType"), node.find(f".//{H}FloatType")
        if st is not None:
            return f"str/{st.get('Cset')}/{st.get('StrSize')}/{st.get('StrPad')}"
        if fl is not None:
            return f"float{8 * int(fl.get('Size'))}/{fl.get('ByteOrder')}"
        return "other"

    def shape(node):
        sp = node.find(f"{H}Dataspace")
        if sp.find(f"{H}ScalarDataspace") is not None:
            return "scalar"
        return [int(d.get("DimSize")) for d in sp.iter(f"{H}Dimension")]

    def attribute(a):
        t, sh = dtype(a), shape(a)
        out = {"type": t, "shape": sh}
        if t.startswith("str"):
            raw = (a.find(f".//{H}DataFromFile").text or "")
            vals = [re.sub(r"\\(.)", r"\1", m) for m in re.findall(r'"((?:[^"\\]|\\.)*)"', raw)]
            out["value"] = vals[0] if sh == "scalar" else vals
            if a.get("Name") == "model_config":
                out["value"] = json.loads(out["value"])
        return out

    def group(g):
        node = {"attrs": {}, "groups": {}, "datasets": {}}
        for ch in g:
            if ch.tag == f"{H}Attribute":
                node["attrs"][ch.get("Name")] = attribute(ch)
            elif ch.tag == f"{H}Group":
                node["groups"][ch.get("Name")] = group(ch)
            elif ch.tag == f"{H}Dataset":
                node["datasets"][ch.get("Name")] = {"type": dtype(ch), "shape": shape(ch)}
        return node

    return group(ET.fromstring(xml).find(f"{H}RootGroup"))
