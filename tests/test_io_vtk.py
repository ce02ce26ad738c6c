"""CPU checks of the spe10 driver's output files (src/spe10.c:203-218,
246-274, 683-695): VTK RectilinearGrid snapshots, ParaView series indexes
and the P_solver.csv log."""
import json
import xml.etree.ElementTree as ET

import numpy as np

from premixfem_amd import io_vtk


def test_vtr_round_trip_and_cell_centre_coordinates(tmp_path):
    rng = np.random.default_rng(0)
    field = rng.standard_normal((3, 4, 5))   # (P, N, M)
    path = tmp_path / "S_w0.vtr"
    io_vtk.write_vtr(path, field, (20.0, 10.0, 2.0), "S_w")
    back, (x, y, z), name = io_vtk.read_vtr(path)
    assert name == "S_w"
    np.testing.assert_array_equal(back, field)
    np.testing.assert_allclose(x, [10.0, 30.0, 50.0, 70.0, 90.0])
    np.testing.assert_allclose(y, [5.0, 15.0, 25.0, 35.0])
    np.testing.assert_allclose(z, [1.0, 3.0, 5.0])
    # the XML header is well formed and names the extent and the array
    raw = open(path, "rb").read()
    head = raw[:raw.index(b"<AppendedData")].decode() + "</VTKFile>"
    root = ET.fromstring(head.replace("</RectilinearGrid>\n  </VTKFile>", "</RectilinearGrid></VTKFile>"))
    grid = root.find("RectilinearGrid")
    assert grid.get("WholeExtent") == "0 4 0 3 0 2"
    names = [d.get("Name") for d in root.iter("DataArray")]
    assert names == ["x", "y", "z", "S_w"]
    assert raw.rstrip().endswith(b"</VTKFile>")


def test_series_index_is_json(tmp_path):
    s = io_vtk.VtrSeries(tmp_path / "S_w.vtr.series")
    s.add("S_w0.vtr", 0.0)
    s.add("S_w5.vtr", 12.3456789012)
    s.close()
    d = io_vtk.read_series(tmp_path / "S_w.vtr.series")
    assert d["file-series-version"] == "1.0"
    assert [f["name"] for f in d["files"]] == ["S_w0.vtr", "S_w5.vtr"]
    assert d["files"][1]["time"] == 12.345678901


def test_csv_log_format(tmp_path):
    """The reference's header (ten names) and row format (nine values)."""
    assert io_vtk.CSV_HEADER.strip().split(",")[-3:] == ["iter_count_base", "S_w_min", "S_w_max"]
    row = io_vtk.csv_row(3, 12.5, 2, 0.001, 50, 0.25, 42, 0.2, 0.61)
    assert row == "3,1.25000E+01,2,0.00100,50,0.25000,42,2.00000E-01,6.10000E-01\n"


def test_impes_output_writes_every_freq_steps(tmp_path):
    mesh = (4, 3, 2)
    out = io_vtk.ImpesOutput(tmp_path, mesh, (20.0, 10.0, 2.0), save_files_freq=2, output_P=True)
    n = 24
    for k in range(3):
        row = dict(step=k, simu_time=1.5 * k, stage=1, DS_max=0.001, S_w_steps=50, its=10 + k,
                   S_w_min=0.2, S_w_max=0.3 + 0.01 * k, t_assemble=0.1, t_setup=0.2, t_solve=0.3)
        out.step(row, np.full(n, 0.2 + k), np.arange(n) + 100.0 * k)
    out.close()
    files = sorted(p.name for p in tmp_path.iterdir())
    assert files == ["P.vtr.series", "P0.vtr", "P2.vtr", "P_solver.csv", "S_w.vtr.series",
                     "S_w0.vtr", "S_w2.vtr"]
    lines = open(tmp_path / "P_solver.csv").read().splitlines()
    assert len(lines) == 4 and lines[2].startswith("1,1.50000E+00,1,")
    p2, _, _ = io_vtk.read_vtr(tmp_path / "P2.vtr")
    np.testing.assert_array_equal(p2.ravel(), np.arange(n) + 200.0)
    assert json.load(open(tmp_path / "S_w.vtr.series"))["files"][1]["time"] == 3.0
