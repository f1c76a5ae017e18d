"""Checkpoint and visualisation files of a simulation (the reference's I/O extensions).

- VTK (ext/WaterLilyWriteVTKExt.jl, ext/WaterLilyReadVTKExt.jl): `vtk_writer`, `save`, `close`
  write one ImageData file (.vti) per call plus a ParaView collection (.pvd); `load_pvd` restarts a
  simulation from the last file of a collection.  The files are VTK XML with raw appended binary
  data (UInt64 block headers, no compression); any VTK reader opens them.  Point data are named
  and laid out as the reference writes them: "Pressure" (Ng…) and "Velocity" with the components
  first (D × Ng…), on the whole ghosted grid (extent 0..N-1 per direction, origin 1, spacing 1).
- Restart files (ext/WaterLilyJLD2Ext.jl): `save_flow`/`load_flow` and `save_meanflow`/
  `load_meanflow` store the same arrays as the reference's JLD2 files (p, u, Δt; P, U, UU, t).
  JLD2 is HDF5 and no HDF5 library is available here, so the container is NumPy's .npz; the
  arrays keep Julia's column-major layout.
"""
import os
import struct
import xml.etree.ElementTree as ET

import ctypes as C
import numpy as np

from . import _abi as A

_VTK_TYPE = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64"}
_NP_TYPE = {"Float32": np.float32, "Float64": np.float64}


def default_attrib():
    """default_attrib() (ext/WaterLilyWriteVTKExt.jl:44-51): Velocity and Pressure."""
    return {"Velocity": lambda sim: sim.flow.u, "Pressure": lambda sim: sim.flow.p}


class VTKWriter:
    def __init__(self, fname, dir_name, attrib, count=0, entries=None):
        self.fname, self.dir_name, self.attrib = fname, dir_name, attrib
        self.count = count
        self.entries = list(entries or [])  # (timestep, file)


def vtk_writer(fname="WaterLily", attrib=None, dir="vtk_data"):
    """vtkWriter(fname; attrib, dir) (ext/WaterLilyWriteVTKExt.jl:30-33)."""
    os.makedirs(dir, exist_ok=True)
    return VTKWriter(fname, dir, attrib or default_attrib())


def _components_first(a, N):
    """components_first (ext/WaterLilyWriteVTKExt.jl:76): u[x,y,(z),i] -> u[i,x,y,(z)] (column-major)."""
    return a if a.shape == N else np.moveaxis(a, -1, 0)


def write_vti(path, arrays, N):
    """One ImageData file: `arrays` maps names to column-major arrays of shape N or (D,)+N."""
    D = len(N)
    ext = []
    for d in range(3):
        ext += [0, (N[d] - 1) if d < D else 0]
    extent = " ".join(str(v) for v in ext)
    blocks, offset = [], 0
    head = ['<?xml version="1.0"?>',
            '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
            f'  <ImageData WholeExtent="{extent}" Origin="1 1 1" Spacing="1 1 1">',
            f'    <Piece Extent="{extent}">', "      <PointData>"]
    for name, a in arrays.items():
        a = np.asarray(a)
        ncomp = 1 if a.shape == N else a.shape[0]
        raw = np.asfortranarray(a).tobytes(order="F")
        head.append(f'        <DataArray type="{_VTK_TYPE[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}" '
                    f'format="appended" offset="{offset}"/>')
        blocks.append(struct.pack("<Q", len(raw)) + raw)
        offset += 8 + len(raw)
    head += ["      </PointData>", "    </Piece>", "  </ImageData>", '  <AppendedData encoding="raw">']
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n   _").encode())
        for b in blocks:
            fh.write(b)
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vti(path):
    """(arrays, N) of a .vti file with raw appended data (as written by write_vti)."""
    data = open(path, "rb").read()
    start = data.index(b'<AppendedData encoding="raw">')
    us = data.index(b"_", start) + 1
    root = ET.fromstring(data[:start].decode() + "</VTKFile>")
    img = root.find("ImageData")
    ext = [int(v) for v in img.get("WholeExtent").split()]
    N = tuple(ext[2 * d + 1] + 1 for d in range(3) if ext[2 * d + 1] > 0)
    out = {}
    for da in img.iter("DataArray"):
        off = us + int(da.get("offset"))
        nbytes = struct.unpack("<Q", data[off:off + 8])[0]
        a = np.frombuffer(data[off + 8:off + 8 + nbytes], dtype=_NP_TYPE[da.get("type")])
        nc = int(da.get("NumberOfComponents", "1"))
        shape = N if nc == 1 else (nc,) + N
        out[da.get("Name")] = a.reshape(shape, order="F").copy(order="F")
    return out, N


def _write_pvd(path, entries):
    lines = ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">',
             "  <Collection>"]
    for t, f in entries:
        lines.append(f'    <DataSet timestep="{t}" part="0" file="{f}"/>')
    lines += ["  </Collection>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_pvd(path):
    root = ET.parse(path).getroot()
    return [(float(d.get("timestep")), d.get("file")) for d in root.find("Collection").iter("DataSet")]


def save(w, sim):
    """save!(w::VTKWriter, sim) (ext/WaterLilyWriteVTKExt.jl:58-66): the attributes at sim_time."""
    from .api import sim_time
    N = tuple(sim.flow.Ng)
    f = os.path.join(w.dir_name, "%s_%06i.vti" % (w.fname, w.count))
    write_vti(f, {name: _components_first(fn(sim), N) for name, fn in w.attrib.items()}, N)
    w.count += 1
    w.entries.append((round(sim_time(sim), 4), f))
    _write_pvd(w.fname + ".pvd", w.entries)  # kept current; close() writes it once more


def close(w):
    """close(w::VTKWriter) (ext/WaterLilyWriteVTKExt.jl:72): writes the collection file."""
    _write_pvd(w.fname + ".pvd", w.entries)


def load_pvd(sim, fname="WaterLily.pvd", attrib=None, pressure="Pressure", velocity="Velocity"):
    """load!(sim; fname) (ext/WaterLilyReadVTKExt.jl:23-44): p and u from the collection's last
    file; Δt[end] = its timestep·L/U, then push!(Δt, CFL(flow)).  Returns a writer that appends."""
    entries = read_pvd(fname)
    arrays, N = read_vti(entries[-1][1])
    if tuple(N) != tuple(sim.flow.Ng):
        raise AssertionError("The dimensions of the simulation do not match the dimensions of the vtk file.")
    sim.flow.set("p", arrays[pressure])
    sim.flow.set("u", np.moveaxis(arrays[velocity], 0, -1))
    T = sim.flow.T
    sim.flow.Δt[-1] = float(T(entries[-1][0] * sim.L / sim.U))
    dt = C.c_double()
    A.check(A.load().wl_flow_cfl(sim.flow.h, C.byref(dt)))
    sim.flow.Δt.append(float(T(dt.value)))
    d = os.path.dirname(entries[0][1])
    return VTKWriter(os.path.splitext(os.path.basename(fname))[0], d, attrib or default_attrib(), len(entries),
                     entries)


def save_flow(fname, sim_or_flow, dir="./"):
    """save!(fname, flow; dir) (ext/WaterLilyJLD2Ext.jl:11-16): p, u and Δt."""
    flow = getattr(sim_or_flow, "flow", sim_or_flow)
    np.savez(os.path.join(dir, fname), p=flow.p, u=flow.u, dt=np.asarray(flow.Δt, dtype=flow.T))


def load_flow(sim_or_flow, fname="WaterLily.npz", dir="./"):
    """load!(flow; fname, dir) (ext/WaterLilyJLD2Ext.jl:38-49)."""
    flow = getattr(sim_or_flow, "flow", sim_or_flow)
    with np.load(os.path.join(dir, fname), allow_pickle=False) as z:
        if tuple(z["p"].shape) != tuple(flow.Ng):
            raise AssertionError("Simulation size does not match the size of the stored simulation.")
        flow.set("p", z["p"])
        flow.set("u", z["u"])
        flow.Δt[:] = [float(v) for v in z["dt"]]


def save_meanflow(fname, mf, dir="./"):
    """save!(fname, meanflow; dir) (ext/WaterLilyJLD2Ext.jl:23-29): P, U, UU and t."""
    kw = dict(P=mf.P, U=mf.U, t=np.asarray(mf.t, dtype=mf.T))
    if mf.uu_stats:
        kw["UU"] = mf.UU
    np.savez(os.path.join(dir, fname), **kw)


def load_meanflow(mf, fname="WaterLilyMean.npz", dir="./"):
    """load!(meanflow; fname, dir) (ext/WaterLilyJLD2Ext.jl:57-70): the stored averages are put
    back on the device."""
    with np.load(os.path.join(dir, fname), allow_pickle=False) as z:
        mf._put("P", z["P"])
        mf._put("U", z["U"])
        if mf.uu_stats and "UU" in z:
            mf._put("UU", z["UU"])
        mf.t = [mf.T(v) for v in z["t"]]
