"""Output files of the spe10 fork's IMPES driver (src/spe10.c:203-218,
246-274, 683-695): the ``P_solver.csv`` step log, ``<name><step>.vtr``
snapshots of S_w and P and the ParaView ``.vtr.series`` index files.

Host-side writers of arrays already copied from the device.  The ``.vtr``
files are VTK XML RectilinearGrid files (the format PETSc's VTK viewer
writes for a DMDA vector, ``save_vec_into_vtr``): one point per cell
centre, the field as Float64 point data, appended raw binary with UInt64
block headers.
"""
from __future__ import annotations

import json
import os

import numpy as np

# The log header of src/spe10.c:207; the reference's rows carry nine of its
# ten columns (iter_count_base is never written), kept so that existing
# post-processing reads these files unchanged.
CSV_HEADER = ("simu_time_ind,simu_time,stage_ind,DS_max,S_w_steps,wall_time,iter_count,"
              "iter_count_base,S_w_min,S_w_max\n")


def csv_row(step, simu_time, stage, DS_max, S_w_steps, wall_time, its, S_w_min, S_w_max):
    """One log row in the reference's format (:255)."""
    return (f"{step:d},{simu_time:.5E},{stage:d},{DS_max:.5f},{S_w_steps:d},{wall_time:.5f},"
            f"{its:d},{S_w_min:.5E},{S_w_max:.5E}\n")


def write_vtr(path, field, spacing, name):
    """Write ``field`` (natural layout, numpy shape (P, N, M)) as a VTK
    RectilinearGrid with points at the cell centres ((i + 1/2) H)."""
    field = np.ascontiguousarray(field, dtype="<f8")
    P, N, M = field.shape
    hx, hy, hz = spacing
    coords = [((np.arange(n) + 0.5) * h).astype("<f8") for n, h in ((M, hx), (N, hy), (P, hz))]
    blocks = coords + [field.ravel()]
    offsets, off = [], 0
    for b in blocks:
        offsets.append(off)
        off += 8 + b.nbytes
    ext = f"0 {M - 1} 0 {N - 1} 0 {P - 1}"
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">\n'
        f'  <RectilinearGrid WholeExtent="{ext}">\n'
        f'    <Piece Extent="{ext}">\n'
        '      <Coordinates>\n'
        f'        <DataArray type="Float64" Name="x" format="appended" offset="{offsets[0]}"/>\n'
        f'        <DataArray type="Float64" Name="y" format="appended" offset="{offsets[1]}"/>\n'
        f'        <DataArray type="Float64" Name="z" format="appended" offset="{offsets[2]}"/>\n'
        '      </Coordinates>\n'
        f'      <PointData Scalars="{name}">\n'
        f'        <DataArray type="Float64" Name="{name}" NumberOfComponents="1" '
        f'format="appended" offset="{offsets[3]}"/>\n'
        '      </PointData>\n'
        '    </Piece>\n'
        '  </RectilinearGrid>\n'
        '  <AppendedData encoding="raw">\n_')
    with open(path, "wb") as f:
        f.write(head.encode())
        for b in blocks:
            f.write(np.uint64(b.nbytes).astype("<u8").tobytes())
            f.write(b.tobytes())
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vtr(path):
    """Read back a file written by :func:`write_vtr`: (field (P, N, M),
    (x, y, z) coordinates, array name)."""
    import re

    raw = open(path, "rb").read()
    cut = raw.index(b"<AppendedData")
    head = raw[:cut].decode()
    start = raw.index(b"_", cut) + 1
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    M, N, P = ext[1] + 1, ext[3] + 1, ext[5] + 1
    arrays = re.findall(r'Name="([^"]+)"[^>]*offset="(\d+)"', head)
    out = {}
    for nm, off in arrays:
        o = start + int(off)
        nbytes = int(np.frombuffer(raw[o:o + 8], dtype="<u8")[0])
        out[nm] = np.frombuffer(raw[o + 8:o + 8 + nbytes], dtype="<f8")
    name = [nm for nm, _ in arrays if nm not in ("x", "y", "z")][0]
    return out[name].reshape(P, N, M), (out["x"], out["y"], out["z"]), name


class VtrSeries:
    """A ParaView ``.vtr.series`` index (:209-216, :263-268), closed by
    :meth:`close`."""

    def __init__(self, path):
        self.f = open(path, "w")
        self.f.write('{\n"file-series-version" : "1.0",\n"files" : [\n')
        self.n = 0

    def add(self, vtr_name, time):
        sep = "" if self.n == 0 else ",\n"
        self.f.write(f'{sep}{{"name" :"{vtr_name}", "time" :{time:.9f}}}')
        self.n += 1

    def close(self):
        if self.f:
            self.f.write("\n]\n}\n")
            self.f.close()
            self.f = None


class ImpesOutput:
    """The driver's outputs in ``out_dir``: ``P_solver.csv`` always; every
    ``save_files_freq`` steps ``S_w<step>.vtr`` (and ``P<step>.vtr`` with
    ``output_P``) with their series files (src/spe10.c:121-143 options)."""

    def __init__(self, out_dir, mesh, spacing, save_files_freq=1, output_P=False):
        os.makedirs(out_dir, exist_ok=True)
        self.dir = out_dir
        self.mesh = tuple(mesh)
        self.spacing = tuple(spacing)
        self.freq = max(1, int(save_files_freq))
        self.output_P = output_P
        self.csv = open(os.path.join(out_dir, "P_solver.csv"), "w")
        self.csv.write(CSV_HEADER)
        self.s_series = VtrSeries(os.path.join(out_dir, "S_w.vtr.series"))
        self.p_series = VtrSeries(os.path.join(out_dir, "P.vtr.series")) if output_P else None

    def step(self, row, S_w, P_):
        """row: one IMPES log row (dict); S_w, P_: natural-order arrays."""
        M, N, Pz = self.mesh
        wall = row["t_assemble"] + row["t_setup"] + row["t_solve"]
        self.csv.write(csv_row(row["step"], row["simu_time"], row["stage"], row["DS_max"],
                               row["S_w_steps"], wall, row["its"], row["S_w_min"],
                               row["S_w_max"]))
        self.csv.flush()
        if row["step"] % self.freq == 0:
            nm = f"S_w{row['step']}.vtr"
            write_vtr(os.path.join(self.dir, nm), np.asarray(S_w).reshape(Pz, N, M), self.spacing,
                      "S_w")
            self.s_series.add(nm, row["simu_time"])
            if self.p_series is not None:
                nm = f"P{row['step']}.vtr"
                write_vtr(os.path.join(self.dir, nm), np.asarray(P_).reshape(Pz, N, M),
                          self.spacing, "P")
                self.p_series.add(nm, row["simu_time"])

    def close(self):
        if self.csv:
            self.csv.close()
            self.csv = None
        self.s_series.close()
        if self.p_series is not None:
            self.p_series.close()


def read_series(path):
    return json.load(open(path))
