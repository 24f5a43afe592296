"""Binary PLY writer / reader for meshes (save_ply of i3d/meshing.py:316-384).

The reference writes through plyfile: one ``vertex`` element with float32 x,
y, z (plus optional attributes) and one ``face`` element with a
``vertex_indices`` list of uchar count and int32 indices, in native
(little-endian) binary.  This writer produces that layout with numpy
structured arrays in one pass instead of a per-element Python loop.
"""
import numpy as np

_PLY_TYPES = {"i1": "char", "u1": "uchar", "i2": "short", "u2": "ushort", "i4": "int",
              "u4": "uint", "f4": "float", "f8": "double"}


def save_ply(verts, faces, filename, vertex_attributes=None):
    """Write verts [V, >=3] and faces [F, 3] to a binary little-endian PLY.

    ``vertex_attributes``: list of (name, numpy type code) for the vertex
    columns after x, y, z (e.g. [("nx", "f4"), ("ny", "f4"), ("nz", "f4")]).
    """
    verts = np.asarray(verts)
    faces = np.asarray(faces)
    dtypes = [("x", "f4"), ("y", "f4"), ("z", "f4")]
    if vertex_attributes is not None:
        dtypes[3:3] = list(vertex_attributes)
    vdt = np.dtype([(n, "<" + t) for n, t in dtypes])
    vrec = np.zeros(len(verts), dtype=vdt)
    for k, (name, _) in enumerate(dtypes):
        vrec[name] = verts[:, k]
    fdt = np.dtype([("n", "u1"), ("idx", "<i4", (3,))])
    frec = np.zeros(len(faces), dtype=fdt)
    frec["n"] = 3
    if len(faces):
        frec["idx"] = faces.reshape(-1, 3)
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {len(verts)}"]
    header += [f"property {_PLY_TYPES[t]} {n}" for n, t in dtypes]
    header += [f"element face {len(faces)}", "property list uchar int vertex_indices", "end_header"]
    with open(filename, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(vrec.tobytes())
        f.write(frec.tobytes())


def read_ply(filename):
    """Read a file written by save_ply (or any binary little-endian PLY with
    scalar vertex properties and uchar/int triangle lists)."""
    with open(filename, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    lines = data[:end].decode("ascii").splitlines()
    if lines[0] != "ply" or lines[1] != "format binary_little_endian 1.0":
        raise ValueError("not a binary little-endian PLY")
    inv = {v: k for k, v in _PLY_TYPES.items()}
    elements, cur = [], None
    for ln in lines[2:]:
        parts = ln.split()
        if parts[0] == "element":
            cur = [parts[1], int(parts[2]), []]
            elements.append(cur)
        elif parts[0] == "property":
            cur[2].append(parts[1:])
    out, off = {}, end
    for name, count, props in elements:
        if props and props[0][0] == "list":
            ct, it = inv[props[0][1]], inv[props[0][2]]
            dt = np.dtype([("n", ct), ("idx", "<" + it, (3,))])
            rec = np.frombuffer(data, dtype=dt, count=count, offset=off)
            if count and not np.all(rec["n"] == 3):
                raise ValueError("only triangle faces are supported")
            out[name] = rec["idx"].copy()
        else:
            dt = np.dtype([(p[1], "<" + inv[p[0]]) for p in props])
            rec = np.frombuffer(data, dtype=dt, count=count, offset=off)
            out[name] = rec
        off += dt.itemsize * count
    return out
