"""Independent Python readings of the reference's configuration and VTI formats,
used to check include/smac_config.hpp (tests only).

* initial_fields(): place_vof (CFD/StaggeredGrid.h:144-163, cell centres strictly
  inside the box) and set_sphere (CFD/SMACSolver.cpp:642-693, alpha = 1 and
  v = vini on both y-faces of cells whose centre is inside the sphere), applied in
  file order on the reference padded layout;
* desc(): context kwargs (grid, per-side bc ids, bc table rows 0..max id with
  unlisted rows no-slip at rest, fluid);
* read_vti(): the ImageData file of FileInOut::output_vti_binary_cellData
  (CFD/FileInOut.cpp:145-255) -- header attributes and the appended raw blocks.
"""
import re
import struct

import numpy as np

from mydemcfd_amd import _abi

BC_TYPES = {"noslip": _abi.BC_NOSLIP, "slip": _abi.BC_SLIP, "inflow": _abi.BC_INFLOW, "outlet": _abi.BC_OUTLET}


def desc(cfg):
    g = cfg["grid"]
    nx, ny, nz = g["cells"]
    nb = max([b["id"] for b in cfg["boundary_conditions"]] + [0]) + 1
    table = [(_abi.BC_NOSLIP, 0.0, 0.0, 0.0)] * nb
    for b in cfg["boundary_conditions"]:
        v = b.get("velocity", [0.0, 0.0, 0.0])
        table[b["id"]] = (BC_TYPES[b["type"]], float(v[0]), float(v[1]), float(v[2]))
    side = [0] * 6
    for f in cfg["boundary_faces"]:
        side[2 * "xyz".index(f["axis"]) + (1 if f["side"] == "max" else 0)] = f["boundary_id"]
    fl = cfg["fluid"]
    return dict(n=(nx, ny, nz), size=tuple(g["size"]), origin=tuple(g["origin"]), periodic=(0, 0, 0),
                side_bcid=tuple(side), bc_table=tuple(table),
                rho=(fl["phase0"]["density"], fl["phase1"]["density"]),
                mu=(fl["phase0"]["viscosity"], fl["phase1"]["viscosity"]), g=tuple(fl["gravity"]))


def initial_fields(cfg):
    g = cfg["grid"]
    nx, ny, nz = g["cells"]
    h = [g["size"][d] / g["cells"][d] for d in range(3)]
    o = g["origin"]
    x = h[0] * np.arange(nx + 2) - 0.5 * h[0] + o[0]
    y = h[1] * np.arange(ny + 2) - 0.5 * h[1] + o[1]
    z = h[2] * np.arange(nz + 2) - 0.5 * h[2] + o[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    inner = np.zeros(Z.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    alpha = np.zeros((nz + 2, ny + 2, nx + 2))
    v = np.zeros((nz + 2, ny + 3, nx + 2))
    for ic in cfg["initial_conditions"]:
        if ic["type"] == "vof_box":
            lo, hi = ic["minimum"], ic["maximum"]
            m = inner & (X < hi[0]) & (X > lo[0]) & (Y < hi[1]) & (Y > lo[1]) & (Z < hi[2]) & (Z > lo[2])
            alpha[m] = ic["alpha"]
        elif ic["type"] == "vof_sphere":
            c, r = ic["center"], ic["radius"]
            dx, dy, dz = X - c[0], Y - c[1], Z - c[2]
            m = inner & (dx * dx + dy * dy + dz * dz < r * r)
            alpha[m] = 1.0
            kk, jj, ii = np.nonzero(m)
            v[kk, jj, ii] = ic["vini"]
            v[kk, jj + 1, ii] = ic["vini"]
    return alpha, v


def read_vti(path):
    """(header attributes, {name: float32 array}) of a VTI file with raw
    appended Float32 blocks behind UInt32 byte counts."""
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b'<AppendedData encoding="raw">\n_')
    text = head.decode()
    attrs = {}
    m = re.search(r'<ImageData WholeExtent="([^"]*)" Origin="([^"]*)" Spacing="([^"]*)"', text)
    attrs["extent"] = [int(t) for t in m.group(1).split()]
    attrs["origin_text"] = m.group(2)
    attrs["spacing_text"] = m.group(3)
    attrs["origin"] = [float(t) for t in m.group(2).split()]
    attrs["spacing"] = [float(t) for t in m.group(3).split()]
    arrays = {}
    for dm in re.finditer(r'<DataArray type="Float32" Name="([^"]+)"(?: NumberOfComponents="(\d)")? '
                          r'format="appended" offset="(\d+)"/>', text):
        name, comps, off = dm.group(1), int(dm.group(2) or 1), int(dm.group(3))
        nb = struct.unpack_from("<I", rest, off)[0]
        arrays[name] = (np.frombuffer(rest, dtype="<f4", count=nb // 4, offset=off + 4), comps)
    attrs["tail_ok"] = rest.endswith(b"\n  </AppendedData>\n</VTKFile>\n")
    return attrs, arrays
