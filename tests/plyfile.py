"""Writes binary little-endian PLY files of splat scenes for the loader tests."""
import numpy as np

from gsr import scenes

NP_TYPES = {"char": "i1", "uchar": "u1", "int8": "i1", "uint8": "u1", "short": "<i2", "ushort": "<u2",
            "int16": "<i2", "uint16": "<u2", "int": "<i4", "uint": "<u4", "int32": "<i4", "uint32": "<u4",
            "float": "<f4", "float32": "<f4", "double": "<f8", "float64": "<f8"}


def property_columns(sc: scenes.Scene):
    cols = {"x": sc.pos[:, 0], "y": sc.pos[:, 1], "z": sc.pos[:, 2]}
    for k in range(3):
        cols[f"scale_{k}"] = sc.lscale[:, k]
        cols[f"f_dc_{k}"] = sc.dc[:, k]
    for k in range(4):
        cols[f"rot_{k}"] = sc.rot[:, k]
    cols["opacity"] = sc.opacity
    for k in range(45):
        cols[f"f_rest_{k}"] = sc.rest[:, k]
    return cols


def write_ply(path, sc: scenes.Scene, order=None, types=None, extra=(), fmt="binary_little_endian",
              drop=(), list_prop=False, count=None, truncate=0, crlf=False, no_end=False):
    cols = property_columns(sc)
    names = list(order) if order else list(cols)
    props = [(n, (types or {}).get(n, "float")) for n in names if n not in drop]
    props += list(extra)
    nl = "\r\n" if crlf else "\n"
    lines = ["ply", f"format {fmt} 1.0", f"element vertex {sc.count if count is None else count}"]
    for n, t in props:
        lines.append(f"property {t} {n}")
    if list_prop:
        lines.append("property list uchar int vertex_indices")
    if not no_end:
        lines.append("end_header")
    head = (nl.join(lines) + nl).encode("ascii")
    dt = np.dtype([(n, NP_TYPES[t]) for n, t in props])
    rec = np.zeros(sc.count, dtype=dt)
    for n, t in props:
        rec[n] = cols[n] if n in cols else 7
    body = rec.tobytes()
    if truncate:
        body = body[:-truncate]
    with open(path, "wb") as f:
        f.write(head + body)
    # the values the loader must produce (after the declared type's rounding)
    expect = {n: rec[n].astype(np.float32) for n, _ in props if n in cols}
    return expect
